"""The three query shapes of the path, as specs over the C-ABI, plus device
generation of their seeded synthetic inputs.

  lineitem   filtered low-cardinality group-by (TPC-H Q1 shape):
             WHERE shipdate <= 2436 GROUP BY flag, status; SUM(quantity),
             SUM(price), SUM(price*(1-discount)), SUM(price*(1-discount)*(1+tax)),
             AVG(quantity), AVG(discount), COUNT(*).
  join       equi join of a filtered build (segment = 1) with a probe table;
             COUNT(*), SUM(value) over the matches and the matched
             (probe_row_id, build_row_id) pairs.
  groups     high-cardinality group-by: COUNT(*), SUM(vi), MIN/MAX(vf),
             SUM(vf) by an int64 key.

AVG is lowered as the reference lowers it (qdb/plan/sql_lowering.cpp:
799-814): SUM and COUNT of the argument in the aggregate, then
decimal_to_f64(SUM) / f64(COUNT) over the result.
"""
from __future__ import annotations

import ctypes as C
from fractions import Fraction

import numpy as np

from . import abi
from .host import (AggSpec, DECIMAL, DATE32, DeviceBuffer, DeviceRowSet, F64, I8, I32, I64, JoinSpec, Program,
                   check, lib)

SEED_LINEITEM = 0xC0
SEED_JOIN = 0xC1
SEED_GROUPS = 0xC2

SHIPDATE_CUTOFF = 2436


def decimal_to_f64(unscaled: int, scale: int) -> float:
    """The nearest double of unscaled / 10^scale, ties to even
    (qdb_decimal_to_f64, qdb/modules/qumirdb.oz:983-1046)."""
    if abs(unscaled) < (1 << 53) and scale <= 22:
        return float(unscaled) / float(10 ** scale)
    return float(Fraction(unscaled, 10 ** scale))


# --------------------------------------------------------------------------
# lineitem
# --------------------------------------------------------------------------

LINEITEM_TYPES = [I8(), I8(), DATE32(), DECIMAL(15, 2), DECIMAL(15, 2), DECIMAL(15, 2), DECIMAL(15, 2)]
FLAG, STATUS, SHIPDATE, QUANTITY, PRICE, DISCOUNT, TAX = range(7)


def lineitem_predicate(types=LINEITEM_TYPES):
    p = Program(types)
    r = p.cmp(abi.CMP_LE, p.col(SHIPDATE), p.const_int(SHIPDATE_CUTOFF))
    return p, r


def lineitem_disc_price(types=LINEITEM_TYPES):
    """price * (1 - discount): DECIMAL(31,4)."""
    p = Program(types)
    r = p.mul(p.col(PRICE), p.sub(p.const_int(1), p.col(DISCOUNT)))
    return p, r


def lineitem_charge(types=LINEITEM_TYPES):
    """price * (1 - discount) * (1 + tax): DECIMAL(38,6)."""
    p = Program(types)
    disc = p.mul(p.col(PRICE), p.sub(p.const_int(1), p.col(DISCOUNT)))
    r = p.mul(disc, p.add(p.const_int(1), p.col(TAX)))
    return p, r


# Aggregates over input + computed columns (7 = disc_price, 8 = charge).
LINEITEM_AGGS = [
    (abi.AGG_SUM, QUANTITY), (abi.AGG_SUM, PRICE), (abi.AGG_SUM, 7), (abi.AGG_SUM, 8),
    (abi.AGG_SUM, QUANTITY), (abi.AGG_COUNT, QUANTITY),   # AVG(quantity)
    (abi.AGG_SUM, DISCOUNT), (abi.AGG_COUNT, DISCOUNT),   # AVG(discount)
    (abi.AGG_COUNT_STAR, -1),
]


def lineitem_fused_spec(output_on_host: bool = False) -> AggSpec:
    """Filter, project and aggregate in one kernel."""
    return AggSpec(types=LINEITEM_TYPES, keys=[FLAG, STATUS], aggs=LINEITEM_AGGS, output_on_host=output_on_host,
                   predicate=lineitem_predicate(), computed=[lineitem_disc_price(), lineitem_charge()])


def lineitem_projected_types():
    return LINEITEM_TYPES + [lineitem_disc_price()[0].result_type(lineitem_disc_price()[1]),
                             lineitem_charge()[0].result_type(lineitem_charge()[1])]


def lineitem_plain_spec() -> AggSpec:
    """The aggregate over the projected batch (filter and project ran as
    their own operators, like the reference's plan filter -> project ->
    aggregate, test/regtest/goldens/tpch/q1.plan)."""
    return AggSpec(types=lineitem_projected_types(), keys=[FLAG, STATUS], aggs=LINEITEM_AGGS)


def lineitem_finalize_on_device(proc, kernels):
    """Finishes the lineitem aggregate in HBM and computes AVG(quantity) and
    AVG(discount) there with the project kernel over the aggregate's output
    (SUM / COUNT converted to DOUBLE, as the reference lowers AVG); returns
    the rows of lineitem_finalize."""
    from .host import ProjectKernel, RowSetView, download_rowset
    types = kernels.out_types
    out = proc.finish_raw()
    try:
        view = RowSetView(out, types)
        avg_q, avg_d = Program(types), Program(types)
        ra = avg_q.avg(6, 7)
        rd = avg_d.avg(8, 9)
        kern = _avg_kernel(types, [(avg_q, ra), (avg_d, rd)])
        bufs, _ = kern(view)
        n = view.rows
        aq = bufs[0].to_numpy(np.float64, n)
        ad = bufs[2 if kern.out_types[0].nullable else 1].to_numpy(np.float64, n)
        cols = download_rowset(out, types)
    finally:
        lib().qdb_gpu_rowset_release(C.byref(out))
    (flag, _), (status, _) = cols[0], cols[1]
    vals = [c[0] for c in cols[2:]]
    rows = []
    for g in range(len(flag)):
        sq, sp, sd, sc, _, _, _, _, cnt = [int(v[g]) for v in vals]
        rows.append((int(flag[g]), int(status[g]), sq, sp, sd, sc, float(aq[g]), float(ad[g]), cnt))
    rows.sort()
    return rows


_AVG_KERNELS: dict = {}


def _avg_kernel(types, outputs):
    """The AVG projection of the lineitem output, compiled once per process."""
    from .host import ProjectKernel
    key = tuple((t.kind, t.width, t.scale, t.nullable) for t in types)
    if key not in _AVG_KERNELS:
        _AVG_KERNELS[key] = ProjectKernel(types, outputs)
    return _AVG_KERNELS[key]


def lineitem_finalize(columns):
    """Rows of (flag, status, sum_qty, sum_price, sum_disc_price, sum_charge,
    avg_qty, avg_disc, count) from an aggregate result, sorted by key."""
    (flag, _), (status, _) = columns[0], columns[1]
    vals = [c[0] for c in columns[2:]]
    rows = []
    for g in range(len(flag)):
        sq, sp, sd, sc, sq2, cq, sdi, cdi, cnt = [int(v[g]) for v in vals]
        rows.append((int(flag[g]), int(status[g]), sq, sp, sd, sc,
                     decimal_to_f64(sq2, 2) / float(cq), decimal_to_f64(sdi, 2) / float(cdi), cnt))
    rows.sort()
    return rows


class LineitemDevice:
    """The lineitem columns generated in HBM."""

    def __init__(self, rows: int, seed: int = SEED_LINEITEM, row0: int = 0):
        self.rows = rows
        widths = [1, 1, 4, 8, 8, 8, 8]
        self.bufs = [DeviceBuffer(max(rows, 1) * w) for w in widths]
        check(lib().qdb_gpu_gen_lineitem(seed, row0, rows, *[b.ptr for b in self.bufs]), "gen_lineitem")
        self.rowset = DeviceRowSet(LINEITEM_TYPES, self.bufs, rows=rows)


# --------------------------------------------------------------------------
# join
# --------------------------------------------------------------------------

BUILD_TYPES = [I64(), I32()]
PROBE_TYPES = [I64(), DECIMAL(15, 2)]
BUILD_KEY, SEGMENT = 0, 1
PROBE_FK, PROBE_VALUE = 0, 1
SEGMENT_FILTER = 1


def join_build_predicate():
    p = Program(BUILD_TYPES)
    r = p.cmp(abi.CMP_EQ, p.col(SEGMENT), p.const_int(SEGMENT_FILTER))
    return p, r


def join_spec(fused: bool = True, table_mode: int = abi.JOIN_TABLE_AUTO) -> JoinSpec:
    return JoinSpec(build_types=BUILD_TYPES, probe_types=PROBE_TYPES, build_keys=[BUILD_KEY], probe_keys=[PROBE_FK],
                    join_type=abi.JOIN_INNER, probe_is_left=True,
                    fused=[(abi.AGG_COUNT_STAR, -1), (abi.AGG_SUM, PROBE_VALUE)] if fused else [],
                    table_mode=table_mode)


class JoinDevice:
    def __init__(self, build_rows: int, probe_rows: int, seed: int = SEED_JOIN):
        self.build_rows, self.probe_rows = build_rows, probe_rows
        self.key = DeviceBuffer(max(build_rows, 1) * 8)
        self.segment = DeviceBuffer(max(build_rows, 1) * 4)
        check(lib().qdb_gpu_gen_join_build(seed, 0, build_rows, self.key.ptr, self.segment.ptr), "gen_join_build")
        self.fk = DeviceBuffer(max(probe_rows, 1) * 8)
        self.value = DeviceBuffer(max(probe_rows, 1) * 8)
        check(lib().qdb_gpu_gen_join_probe(seed, 0, probe_rows, build_rows, self.fk.ptr, self.value.ptr), "gen_join_probe")
        self.build = DeviceRowSet(BUILD_TYPES, [self.key, self.segment], rows=build_rows)
        self.probe = DeviceRowSet(PROBE_TYPES, [self.fk, self.value], rows=probe_rows)


# --------------------------------------------------------------------------
# groups
# --------------------------------------------------------------------------

GROUPS_TYPES = [I64(), I64(), F64()]
GKEY, GVI, GVF = range(3)
GROUPS_AGGS = [(abi.AGG_COUNT_STAR, -1), (abi.AGG_SUM, GVI), (abi.AGG_MIN, GVF), (abi.AGG_MAX, GVF), (abi.AGG_SUM, GVF)]


def groups_spec() -> AggSpec:
    return AggSpec(types=GROUPS_TYPES, keys=[GKEY], aggs=GROUPS_AGGS)


def zipf1_cdf(n: int) -> np.ndarray:
    """Zipf(s=1) CDF over ranks 1..n summed in rank order (the same
    arithmetic as qdb_gen_zipf1_cdf)."""
    h = np.cumsum(1.0 / np.arange(1, n + 1, dtype=np.float64))
    # numpy's cumsum adds in order, as the C loop does.
    cdf = h / h[-1]
    cdf[-1] = 1.0
    return cdf


def zipf_cdf(n: int, s: float) -> np.ndarray:
    """Zipf(s) CDF over ranks 1..n (s = 0: uniform)."""
    w = np.arange(1, n + 1, dtype=np.float64) ** (-s)
    h = np.cumsum(w)
    cdf = h / h[-1]
    cdf[-1] = 1.0
    return cdf


class GroupsDevice:
    def __init__(self, rows: int, distinct: int, seed: int = SEED_GROUPS, cdf: np.ndarray | None = None):
        self.rows, self.distinct = rows, distinct
        cdf = zipf1_cdf(distinct) if cdf is None else cdf
        self.cdf = DeviceBuffer.from_numpy(cdf)
        self.key = DeviceBuffer(max(rows, 1) * 8)
        self.vi = DeviceBuffer(max(rows, 1) * 8)
        self.vf = DeviceBuffer(max(rows, 1) * 8)
        check(lib().qdb_gpu_gen_zipf_groups(seed, 0, rows, self.cdf.ptr, distinct, self.key.ptr, self.vi.ptr, self.vf.ptr),
              "gen_groups")
        self.rowset = DeviceRowSet(GROUPS_TYPES, [self.key, self.vi, self.vf], rows=rows)


# --------------------------------------------------------------------------
# batched form: many group-by instances of one aggregate
# --------------------------------------------------------------------------

SEED_BATCHED = 0xC3


def batched_instances(n: int = 512, seed: int = SEED_BATCHED, rows_lo: int = 10_000, rows_hi: int = 20_000_000,
                      card_lo: int = 4, card_hi: int = 1_000_000):
    """(rows, cardinality, zipf exponent) per instance: rows and cardinality
    log-uniform, keys uniform or Zipf(1.1) over the cardinality."""
    rng = np.random.default_rng(seed)
    rows = np.exp(rng.uniform(np.log(rows_lo), np.log(rows_hi), n)).astype(np.int64)
    card = np.exp(rng.uniform(np.log(card_lo), np.log(card_hi), n)).astype(np.int64)
    s = np.where(rng.random(n) < 0.5, 0.0, 1.1)
    return [(int(r), int(c), float(z)) for r, c, z in zip(rows, card, s)]


class BatchedDevice:
    """Every instance's group-by columns (int64 key, int64 and f64 values,
    as the high-cardinality shape) generated in HBM."""

    def __init__(self, instances, seed: int = SEED_BATCHED):
        self.instances = list(instances)
        self.devs = []
        for i, (rows, card, s) in enumerate(self.instances):
            self.devs.append(GroupsDevice(rows, card, seed=seed * 1000003 + i, cdf=zipf_cdf(card, s)))


# --------------------------------------------------------------------------
# batched form: join instances
# --------------------------------------------------------------------------


def batched_join_instances(n: int = 256, seed: int = SEED_BATCHED + 1, build_lo: int = 1_000, build_hi: int = 2_000_000,
                           rows_lo: int = 10_000, rows_hi: int = 20_000_000):
    """(build rows, probe rows) per join instance, both log-uniform: build
    keys 1..B shuffled (unique), probe foreign keys uniform over 1..B."""
    rng = np.random.default_rng(seed)
    b = np.exp(rng.uniform(np.log(build_lo), np.log(build_hi), n)).astype(np.int64)
    p = np.exp(rng.uniform(np.log(rows_lo), np.log(rows_hi), n)).astype(np.int64)
    return [(int(x), int(y)) for x, y in zip(b, p)]


def batched_join_spec() -> JoinSpec:
    """The join of every batched join instance: build (key i64, segment i32)
    unfiltered, probe (fk i64, value DECIMAL(15,2)); COUNT(*) and SUM(value)
    fused, and the matched pairs."""
    return join_spec(fused=True)


class BatchedJoinDevice:
    """Every join instance's build and probe columns generated in HBM."""

    def __init__(self, instances, seed: int = SEED_BATCHED + 1):
        self.instances = list(instances)
        self.devs = [JoinDevice(b, p, seed=seed * 1000003 + i) for i, (b, p) in enumerate(self.instances)]

