"""Parity of the HIP path with the reference's goldens and the CPU oracle on
the same seeded inputs. Integer, decimal, key and pair results are compared
bit-exactly; f64 sums exactly against the oracle's correctly rounded sum and
within 1e-12 relative against its Kahan-Babuska-Neumaier sum."""
import ctypes as C
import math

import numpy as np
import pytest

import cases
from cases import (AGG_COUNT, AGG_COUNT_STAR, AGG_MAX, AGG_MIN, AGG_SUM, AggCase, Batch, GpuBackend, JoinCase,
                   OracleBackend, random_agg_case)
from helpers import DECIMAL, FLOAT, SIGNED, T, result_rows, sort_key

from oracle import oracle as O

pytestmark = pytest.mark.gpu

F64_REL_TOL = 1e-12


@pytest.fixture(scope="module")
def gpu_backend():
    return GpuBackend()


def same(got, want):
    assert len(got) == len(want), f"{len(got)} rows vs {len(want)}\n{got[:5]}\n{want[:5]}"
    for g, w in zip(got, want):
        assert g == w, f"{g} vs {w}"


def test_device_generators_match_host(gpu_backend):
    """Every full-size parity test feeds device-generated columns to the HIP
    path and host-generated ones to the oracle: the two generators must agree
    bit for bit (row offsets included)."""
    from qumirdb_amd import workloads as W
    n, row0 = 300_000, 12_345
    dev = W.LineitemDevice(n, row0=row0)
    host = O.gen_lineitem(W.SEED_LINEITEM, row0, n)
    for buf, h in zip(dev.bufs, host):
        assert np.array_equal(buf.to_numpy(h.dtype, n), h)
    jd = W.JoinDevice(200_000, 300_000)
    key, seg = O.gen_join_build(W.SEED_JOIN, 0, 200_000, 200_000)
    fk, val = O.gen_join_probe(W.SEED_JOIN, 0, 300_000, 200_000)
    assert np.array_equal(jd.key.to_numpy(np.int64, 200_000), key)
    assert np.array_equal(jd.segment.to_numpy(np.int32, 200_000), seg)
    assert np.array_equal(jd.fk.to_numpy(np.int64, 300_000), fk)
    assert np.array_equal(jd.value.to_numpy(np.int64, 300_000), val)
    gd = W.GroupsDevice(300_000, 100_000)
    gk, gvi, gvf = O.gen_groups(W.SEED_GROUPS, 0, 300_000, W.zipf1_cdf(100_000))
    assert np.array_equal(gd.key.to_numpy(np.int64, 300_000), gk)
    assert np.array_equal(gd.vi.to_numpy(np.int64, 300_000), gvi)
    assert np.array_equal(gd.vf.to_numpy(np.float64, 300_000).view(np.int64), gvf.view(np.int64))


@pytest.fixture(scope="module")
def protocol_backend(gpu_backend):
    b = cases.ProtocolBackend()
    yield b
    b.uninstall()


@pytest.mark.parametrize("name", sorted(cases.AGG_GOLDENS))
def test_aggregate_goldens_through_reference_protocol(protocol_backend, name):
    """Every aggregate golden through the reference's unchanged caller
    protocol: Init / Update, Size read before agg_finish_rowset and equal to
    the finalized row count, the kernel-owned output adopted from a task heap
    of owner-headed blocks and released by DestroyKernelOwnedRowSet with
    nothing left behind (AVG columns through the project kernel)."""
    got, want = cases.AGG_GOLDENS[name](protocol_backend)
    same(got, want)


@pytest.mark.parametrize("name", sorted(cases.AGG_GOLDENS))
def test_aggregate_goldens(gpu_backend, name):
    got, want = cases.AGG_GOLDENS[name](gpu_backend)
    same(got, want)


@pytest.mark.parametrize("name", sorted(cases.JOIN_GOLDENS))
def test_join_goldens(gpu_backend, name):
    got, want = cases.JOIN_GOLDENS[name](gpu_backend)
    same(got, want)


@pytest.mark.parametrize("name", sorted(cases.PROJECT_GOLDENS))
def test_project_goldens(gpu_backend, name):
    got, want = cases.PROJECT_GOLDENS[name](gpu_backend)
    same(got, want)


@pytest.mark.parametrize("seed,n,capacity", [(0, 20000, 8), (1, 50000, 8), (3, 200000, 1 << 14)])
def test_aggregate_random_vs_oracle(gpu_backend, seed, n, capacity):
    """Nullable multi-column keys (one 16-byte decimal argument), every
    reducer; a small initial capacity forces the HBM table to grow and the
    spilled rows and LDS groups to be replayed."""
    case, b = random_agg_case(seed, n)
    got = sorted(result_rows(gpu_backend.aggregate(case, [b], capacity=capacity)), key=sort_key)
    want = sorted(result_rows(OracleBackend(fsum_mode=2).aggregate(case, [b])), key=sort_key)
    same(got, want)


def test_generated_kernel_equals_generic(gpu_backend, monkeypatch):
    """The update kernel generated for an aggregate and the precompiled
    generic kernel give identical results (f64 sums included: both sum
    exactly). QDB_GPU_JIT=0 selects the generic kernel for aggregates
    compiled while it is set."""
    import qumirdb_amd as q
    from qumirdb_amd import workloads as W
    case, b = random_agg_case(7, 40000)
    got_jit = sorted(result_rows(gpu_backend.aggregate(case, [b], capacity=8)), key=sort_key)
    assert q.lib().qdb_gpu_jit_stats(None) > 0, "no kernel was generated"
    monkeypatch.setenv("QDB_GPU_JIT", "0")
    got_generic = sorted(result_rows(gpu_backend.aggregate(case, [b], capacity=8)), key=sort_key)
    same(got_jit, got_generic)
    # the fused lineitem aggregate (predicate and computed i128 arguments)
    outs = []
    for jit in ("1", "0"):
        monkeypatch.setenv("QDB_GPU_JIT", jit)
        dev = W.LineitemDevice(200_000)
        kern = q.AggregateKernels(W.lineitem_fused_spec())
        p = q.TAggregateProcessor(kern, initial_capacity=8)
        p.add(dev.rowset)
        outs.append(W.lineitem_finalize_on_device(p, kern))
        p.close()
    assert outs[0] == outs[1]


@pytest.mark.parametrize("own_path", [False, True], ids=["queued", "own_streams"])
def test_batched_update_equals_plain(gpu_backend, monkeypatch, own_path):
    """Many tables of one aggregate updated in one persistent launch (or,
    own_streams: every instance forced onto the bucketed path, which runs
    each instance on its own, spread over the library stream and two helper
    streams), and
    finished with one synchronization, give exactly the results of one
    update and one finish per table: instances from empty to 300K rows, few
    to many groups, nullable multi-column keys, every reducer; the small
    initial capacities make several tables grow and replay their leftover
    rows inside the batched call."""
    import qumirdb_amd as q
    from cases import random_agg_case
    sizes = [0, 1, 7, 100, 2000, 5000, 33000, 65536, 100000, 300000, 50, 40000]
    insts = [random_agg_case(11 + i, n) for i, n in enumerate(sizes)]
    case = insts[0][0]
    if own_path:
        # The bucketed path stages every argument: no MIN/MAX of 16-byte decimals.
        case = cases.AggCase(case.types, case.keys,
                             [(f, c) for f, c in case.aggs if not (f in (AGG_MIN, AGG_MAX) and case.types[c].width == 16)])
    kern = q.AggregateKernels.__new__(q.AggregateKernels)
    kern.spec = None
    kern.c_spec = gpu_backend.cast(case.to_c())
    kern.k = q.abi.QdbAggregateKernels()
    q.check(q.lib().qdb_gpu_compile_aggregate(C.byref(kern.c_spec), C.byref(kern.k)), "compile_aggregate")
    kern.out_types = [gpu_backend.H.ValueType(t.Kind, t.Width, t.Scale, bool(t.Nullable))
                      for t in kern.k.OutputTypes[: kern.k.NumOutputColumns]]
    devs = [gpu_backend.dev(b) for _, b in insts]
    caps = [8 if i % 3 == 0 else 1 << 16 for i in range(len(insts))]
    plain = []
    for d, cap in zip(devs, caps):
        p = q.TAggregateProcessor(kern, initial_capacity=cap)
        p.add(d)
        plain.append(sorted(result_rows(p.finish()), key=sort_key))
        p.close()
    procs = [q.TAggregateProcessor(kern, initial_capacity=cap) for cap in caps]
    if own_path:
        monkeypatch.setenv("QDB_GPU_TUNE_FORCE", "1,0,0,0,512")
    q.TAggregateProcessor.add_batched(procs, devs)
    # a second batched round on the same tables accumulates
    q.TAggregateProcessor.add_batched(procs[:3], devs[:3])
    monkeypatch.delenv("QDB_GPU_TUNE_FORCE", raising=False)
    finished = q.TAggregateProcessor.finish_batched(procs)
    for i, p in enumerate(procs):
        got = sorted(result_rows(finished[i]), key=sort_key)
        p.close()
        if i < 3:
            _, b = insts[i]
            want = sorted(result_rows(OracleBackend(fsum_mode=2).aggregate(case, [b, b])), key=sort_key)
        else:
            want = plain[i]
        same(got, want)
    q.lib().qdb_gpu_release_slot(q.abi.SLOT_AGGREGATE, kern.k.Slot)


def _bucketed_case(case):
    """The aggregate without i128 MIN/MAX (those words are applied under a
    lock and keep an update on the plain path)."""
    keep = [(a, c) for a, c in case.aggs if not (a in (AGG_MIN, AGG_MAX) and c == 3)]
    return AggCase(case.types, case.keys, keep)


def test_bucketed_aggregate_random_vs_oracle(gpu_backend, monkeypatch):
    """The bucketed path (rows staged and partitioned by key-hash bucket,
    each bucket aggregated in LDS, merged into the HBM table), taken here
    from 2^20 rows (QDB_GPU_BUCKETED_MIN_ROWS) into a table of at least 2^20
    slots: nullable two-column keys, an i128 SUM, a selection, COUNT(*),
    COUNT, SUM/MIN/MAX of i64, f64 and decimals; then a second batch into
    the same table through the same path."""
    monkeypatch.setenv("QDB_GPU_BUCKETED_MIN_ROWS", str(1 << 20))
    case, b = random_agg_case(21, 1_200_000)
    case = _bucketed_case(case)
    got = sorted(result_rows(gpu_backend.aggregate(case, [b, b], capacity=1 << 21)), key=sort_key)
    want = sorted(result_rows(OracleBackend(fsum_mode=2).aggregate(case, [b, b])), key=sort_key)
    same(got, want)


def test_bucketed_aggregate_skewed_keys_vs_oracle(gpu_backend, monkeypatch):
    """The bucketed path with hot keys (40% and 20% of the rows on two keys,
    the rest over 200K keys): whole waves of a bucket's rows share a key
    and are combined across the wave before they reach the LDS record."""
    monkeypatch.setenv("QDB_GPU_BUCKETED_MIN_ROWS", str(1 << 20))
    rng = np.random.default_rng(77)
    case, b = random_agg_case(22, 1_100_000)
    case = _bucketed_case(case)
    n = len(b.columns[1])
    u = rng.random(n)
    k2 = np.where(u < 0.4, 7, np.where(u < 0.6, -3, rng.integers(0, 200_000, n))).astype(np.int64)
    k1 = np.where(u < 0.6, 1, b.columns[0]).astype(np.int32)
    k1v = np.where(u < 0.6, True, b.valid[0])
    b = Batch(b.types, [k1, k2] + list(b.columns[2:]), [k1v] + list(b.valid[1:]), b.selection)
    got = sorted(result_rows(gpu_backend.aggregate(case, [b, b], capacity=1 << 21)), key=sort_key)
    want = sorted(result_rows(OracleBackend(fsum_mode=2).aggregate(case, [b, b])), key=sort_key)
    same(got, want)


def test_aggregate_batches_accumulate(gpu_backend):
    case, b = random_agg_case(5, 30000)
    half = 15000
    b1 = Batch(b.types, [c[:half] for c in b.columns], [None if v is None else v[:half] for v in b.valid], b.selection[:half])
    b2 = Batch(b.types, [c[half:] for c in b.columns], [None if v is None else v[half:] for v in b.valid], b.selection[half:])
    got = sorted(result_rows(gpu_backend.aggregate(case, [b1, b2])), key=sort_key)
    want = sorted(result_rows(OracleBackend(fsum_mode=2).aggregate(case, [b])), key=sort_key)
    same(got, want)


@pytest.mark.parametrize("table_mode", [0, 1])
@pytest.mark.parametrize("seed", range(2))
def test_join_random_vs_oracle(gpu_backend, seed, table_mode):
    rng = np.random.default_rng(seed)
    nb, npr = 5000 + seed * 40000, 200000
    types_b = [T(SIGNED, 8, nullable=True), T(SIGNED, 4), T(DECIMAL, 16, 2, True, 30)]
    types_p = [T(SIGNED, 8, nullable=True), T(DECIMAL, 8, 2, False, 15)]
    bk = rng.integers(0, nb // 2, nb).astype(np.int64)
    bv = rng.random(nb) > 0.05
    seg = rng.integers(0, 5, nb).astype(np.int32)
    pay = [int(x) * 10 ** 18 for x in rng.integers(-1000, 1000, nb)]
    payv = rng.random(nb) > 0.1
    pk = rng.integers(0, nb, npr).astype(np.int64)
    pv = rng.random(npr) > 0.05
    val = rng.integers(90000, 10500000, npr).astype(np.int64)
    bsel = (seg != 0).astype(np.uint8)
    for jt in (cases.JOIN_INNER, cases.JOIN_LEFT_OUTER):
        case = JoinCase(types_b, types_p, [0], [0], join_type=jt, probe_is_left=True, table_mode=table_mode)
        build = Batch(types_b, [bk, seg, pay], [bv, None, payv], bsel)
        probe = Batch(types_p, [pk, val], [pv, None])
        gp, grows = gpu_backend.join(case, [build], probe)
        op, orows = OracleBackend().join(case, [build], probe)
        assert np.array_equal(gp[np.lexsort((gp[:, 1], gp[:, 0]))], op[np.lexsort((op[:, 1], op[:, 0]))])
        assert sorted(grows, key=sort_key) == sorted(orows, key=sort_key)


def test_join_duplicate_build_sizes_pairs_by_key_multiplicity(gpu_backend):
    """A one-to-many build (200K rows over 50K keys: most keys 4 rows, some
    9) against a 3M-row probe: the pair buffer is sized by the most rows of
    one key, not by the build's duplicate total (which would ask for ~4.5e11
    pairs), and the pairs (about 12M) equal the oracle's."""
    rng = np.random.default_rng(11)
    nkeys, nb, npr = 50_000, 200_000, 3_000_000
    bk = np.concatenate([np.repeat(np.arange(nkeys, dtype=np.int64), 4), np.repeat(np.arange(5000, dtype=np.int64), 5)])
    bk = bk[rng.permutation(len(bk))][:nb]
    pk = rng.integers(-1000, nkeys + 1000, npr).astype(np.int64)
    types_b = [T(SIGNED, 8)]
    types_p = [T(SIGNED, 8), T(SIGNED, 8)]
    pv = np.arange(npr, dtype=np.int64)
    case = JoinCase(types_b, types_p, [0], [0], join_type=cases.JOIN_INNER, probe_is_left=True)
    kern, proc = gpu_backend.join_processor(case)
    proc.append(gpu_backend.dev(Batch(types_b, [bk])))
    proc.build()
    got = proc.probe_pairs(gpu_backend.dev(Batch(types_p, [pk, pv])))
    cap = int(q_pairs_capacity(gpu_backend, proc))
    per_key = np.bincount(bk, minlength=nkeys + 1)
    assert per_key.max() == 9
    assert cap <= max(npr * 9, 1), "pair buffer sized by the most rows of one key"
    oj = O.OracleJoin(case.to_c())
    oj.append(O.HostRowSet(types_b, [bk]))
    oj.build()
    want = oj.probe(O.HostRowSet(types_p, [pk, pv]))
    assert len(got) == len(want) > 10_000_000
    assert np.array_equal(got[np.lexsort((got[:, 1], got[:, 0]))], want[np.lexsort((want[:, 1], want[:, 0]))])
    proc.close()


def q_pairs_capacity(gpu_backend, proc):
    return gpu_backend.q.lib().qdb_gpu_join_processor_pairs(proc.p).contents.Capacity


@pytest.mark.parametrize("table_mode", [0, 1])
@pytest.mark.parametrize("join_type", [cases.JOIN_INNER, cases.JOIN_LEFT_OUTER])
def test_join_distinct_dense_keys_vs_oracle(gpu_backend, join_type, table_mode):
    """Distinct integer build keys over a small range (the dense direct table
    serves them in mode 0) with NULL keys, a build selection, negative keys,
    probe keys outside the range, and fused COUNT/SUM/MIN over a build
    column: the same results as the oracle and as the hash table."""
    rng = np.random.default_rng(7 + join_type)
    nb, npr = 30000, 300000
    types_b = [T(SIGNED, 8, nullable=True), T(SIGNED, 8)]
    types_p = [T(SIGNED, 8, nullable=True), T(SIGNED, 8)]
    bk = rng.permutation(nb).astype(np.int64) - 5000
    bv = rng.random(nb) > 0.02
    bw = rng.integers(-10 ** 12, 10 ** 12, nb).astype(np.int64)
    bsel = (rng.random(nb) > 0.3).astype(np.uint8)
    pk = rng.integers(-9000, nb, npr).astype(np.int64)
    pv = rng.random(npr) > 0.03
    pval = rng.integers(-10 ** 9, 10 ** 9, npr).astype(np.int64)
    case = JoinCase(types_b, types_p, [0], [0], join_type=join_type, probe_is_left=True, table_mode=table_mode,
                    fused=[(AGG_COUNT_STAR, -1), (AGG_SUM, -3), (AGG_MIN, -3), (AGG_SUM, 1)])
    build = Batch(types_b, [bk, bw], [bv, None], bsel)
    probe = Batch(types_p, [pk, pval], [pv, None])
    kern, proc = gpu_backend.join_processor(case)
    proc.append(gpu_backend.dev(build))
    proc.build()
    pairs = proc.probe_pairs(gpu_backend.dev(probe))
    fused = proc.fused_results()
    oj = O.OracleJoin(case.to_c())
    oj.append(OracleBackend().host(build))
    oj.build()
    opairs = oj.probe(OracleBackend().host(probe))
    srt = lambda p: p[np.lexsort((p[:, 1], p[:, 0]))]
    assert np.array_equal(srt(pairs), srt(opairs))
    # The fused aggregates equal an aggregate over the join's output rows:
    # an unmatched probe row of the outer join counts, its build side NULL.
    m = opairs[opairs[:, 1] >= 0]
    prow = (opairs[:, 0] + (1 << 32)).astype(np.int64)
    assert fused[0]["value"] == len(opairs)
    assert fused[1]["value"] == int(bw[m[:, 1]].sum(dtype=np.int64))
    assert fused[2]["value"] == int(bw[m[:, 1]].min())
    assert fused[3]["value"] == int(pval[prow].sum(dtype=np.int64))


@pytest.mark.parametrize("nb", [600_000, 2_100_000])
def test_join_dense_packed_ordinals_equal_unpacked(gpu_backend, monkeypatch, nb):
    """A dense key range above 512K keys gets its ordinals packed at
    ceil(log2(appended rows + 1)) bits (20 and 22 bits here, entries
    straddling word boundaries; a build selection leaves rows out, so the
    ordinals run past the kept row count) for the lean probe: pairs and
    fused results equal the
    unpacked table's (QDB_GPU_JOIN_PACK=0) and the oracle's, with probe keys
    below, inside and above the range."""
    rng = np.random.default_rng(nb)
    types_b = [T(SIGNED, 8), T(SIGNED, 4)]
    types_p = [T(SIGNED, 8), T(SIGNED, 8)]
    # Non-negative keys: the dense table spans the keys' unsigned range.
    bk = rng.permutation(nb).astype(np.int64) + 1234
    bsel = (rng.random(nb) > 0.4).astype(np.uint8)
    npr = 2_000_003
    pk = rng.integers(-2000, nb + 4000, npr).astype(np.int64)
    pval = rng.integers(-10 ** 12, 10 ** 12, npr).astype(np.int64)
    case = JoinCase(types_b, types_p, [0], [0], probe_is_left=True, fused=[(AGG_COUNT_STAR, -1), (AGG_SUM, 1)])
    build = Batch(types_b, [bk, rng.integers(0, 5, nb).astype(np.int32)], None, bsel)
    probe = Batch(types_p, [pk, pval])
    got = {}
    for pack in ("1", "0"):
        monkeypatch.setenv("QDB_GPU_JOIN_PACK", pack)
        kern, proc = gpu_backend.join_processor(case)
        proc.append(gpu_backend.dev(build))
        proc.build()
        info = proc.table_info()
        bits = int(nb).bit_length()   # ordinals index the appended rows
        assert info["dense"] and info["appended"] == nb and info["packed_bits"] == (bits if pack == "1" else 0), info
        pairs = proc.probe_pairs(gpu_backend.dev(probe))
        got[pack] = (pairs[np.lexsort((pairs[:, 1], pairs[:, 0]))], [f["value"] for f in proc.fused_results()])
        proc.close()
    oj = O.OracleJoin(case.to_c())
    oj.append(OracleBackend().host(build))
    oj.build()
    opairs = oj.probe(OracleBackend().host(probe))
    opairs = opairs[np.lexsort((opairs[:, 1], opairs[:, 0]))]
    assert np.array_equal(got["1"][0], opairs) and np.array_equal(got["0"][0], opairs)
    assert got["1"][1] == got["0"][1] == [len(opairs), int(pval[(opairs[:, 0] + (1 << 32)).astype(np.int64)].sum())]


@pytest.mark.parametrize("probe_is_left", [True, False])
@pytest.mark.parametrize("fused", ["count", "sum", "minmax"])
@pytest.mark.parametrize("npr", [1024, 300_007])
def test_join_dense_lean_equals_general(gpu_backend, monkeypatch, probe_is_left, fused, npr):
    """The dense probe's lean form (non-nullable 8-byte key, no selection,
    inner join, fused aggregates over one probe column) against its general
    form (QDB_GPU_JOIN_LEAN=0) and the oracle: both pair orders, with and
    without MIN/MAX, a probe of exactly one wave tile and a ragged one."""
    rng = np.random.default_rng(npr + (3 if probe_is_left else 5))
    nb = 20000
    types_b = [T(SIGNED, 8), T(SIGNED, 4)]
    types_p = [T(SIGNED, 8), T(SIGNED, 8)]
    bk = rng.permutation(nb).astype(np.int64) + 17
    bsel = (rng.random(nb) > 0.6).astype(np.uint8)
    pk = rng.integers(0, nb + 40, npr).astype(np.int64)
    pval = rng.integers(-10 ** 14, 10 ** 14, npr).astype(np.int64)
    specs = {"count": [(AGG_COUNT_STAR, -1)], "sum": [(AGG_COUNT_STAR, -1), (AGG_SUM, 1)],
             "minmax": [(AGG_COUNT_STAR, -1), (AGG_SUM, 1), (AGG_MIN, 1), (AGG_MAX, 1)]}[fused]
    case = JoinCase(types_b, types_p, [0], [0], probe_is_left=probe_is_left, fused=specs)
    build = Batch(types_b, [bk, rng.integers(0, 5, nb).astype(np.int32)], None, bsel)
    probe = Batch(types_p, [pk, pval])
    got = {}
    for lean in ("1", "0"):
        monkeypatch.setenv("QDB_GPU_JOIN_LEAN", lean)
        kern, proc = gpu_backend.join_processor(case)
        proc.append(gpu_backend.dev(build))
        proc.build()
        pairs = proc.probe_pairs(gpu_backend.dev(probe))
        got[lean] = (pairs[np.lexsort((pairs[:, 1], pairs[:, 0]))], [(f["value"], f.get("count")) for f in proc.fused_results()])
        proc.close()
    oj = O.OracleJoin(case.to_c())
    oj.append(OracleBackend().host(build))
    oj.build()
    opairs = oj.probe(OracleBackend().host(probe))
    opairs = opairs[np.lexsort((opairs[:, 1], opairs[:, 0]))]
    assert np.array_equal(got["1"][0], opairs) and np.array_equal(got["0"][0], opairs)
    assert got["1"][1] == got["0"][1]
    prow = (opairs[:, 0 if probe_is_left else 1] + (1 << 32)).astype(np.int64)
    assert got["1"][1][0][0] == len(opairs)
    if fused != "count":
        assert got["1"][1][1][0] == int(sum(int(v) for v in pval[prow]))
    if fused == "minmax":
        assert got["1"][1][2][0] == int(pval[prow].min()) and got["1"][1][3][0] == int(pval[prow].max())


@pytest.mark.parametrize("extremes", ["none", "max", "both"])
@pytest.mark.parametrize("table_mode", [0, 1])
def test_join_hash_build_key_extremes(gpu_backend, extremes, table_mode):
    """The hash build's slot claim with one-word keys takes an empty key
    word outside the build keys' range (above the largest, else below the
    smallest) and falls back to the locked protocol when the keys span the
    whole range: keys with INT64_MAX and/or INT64_MIN, duplicates (chains),
    probed with the same extremes, against the oracle."""
    rng = np.random.default_rng(31 + len(extremes))
    types = [T(SIGNED, 8)]
    bk = rng.integers(-2000, 2000, 6000).astype(np.int64)
    ext = {"none": [], "max": [np.iinfo(np.int64).max], "both": [np.iinfo(np.int64).max, np.iinfo(np.int64).min]}[extremes]
    bk = np.concatenate([bk, np.array(ext * 3, np.int64)])
    pk = np.concatenate([rng.integers(-2500, 2500, 50000).astype(np.int64), np.array(ext * 5 + [0, -1], np.int64)])
    case = JoinCase(types, types, [0], [0], table_mode=table_mode)
    pairs, _ = gpu_backend.join(case, [Batch(types, [bk])], Batch(types, [pk]))
    oj = O.OracleJoin(case.to_c())
    oj.append(OracleBackend().host(Batch(types, [bk])))
    oj.build()
    opairs = oj.probe(OracleBackend().host(Batch(types, [pk])))
    srt = lambda p: p[np.lexsort((p[:, 1], p[:, 0]))]
    assert len(pairs) == len(opairs) > 0
    assert np.array_equal(srt(pairs), srt(opairs))


@pytest.mark.parametrize("table_mode", [0, 1])
def test_join_empty_probe_adds_nothing_to_fused(gpu_backend, table_mode):
    """An empty probe batch, before and after a real one, leaves the fused
    COUNT/SUM/MIN equal to the real batch's alone (an empty batch runs no
    kernel and has no fused words to add)."""
    rng = np.random.default_rng(23)
    types_b = [T(SIGNED, 8)]
    types_p = [T(SIGNED, 8), T(SIGNED, 8)]
    bk = rng.permutation(5000).astype(np.int64)
    pk = rng.integers(0, 6000, 40000).astype(np.int64)
    pval = rng.integers(-10 ** 9, 10 ** 9, 40000).astype(np.int64)
    case = JoinCase(types_b, types_p, [0], [0], table_mode=table_mode,
                    fused=[(AGG_COUNT_STAR, -1), (AGG_SUM, 1), (AGG_MIN, 1)])
    kern, proc = gpu_backend.join_processor(case)
    proc.append(gpu_backend.dev(Batch(types_b, [bk])))
    proc.build()
    empty = Batch(types_p, [np.zeros(0, np.int64), np.zeros(0, np.int64)])
    assert len(proc.probe_pairs(gpu_backend.dev(empty))) == 0
    pairs = proc.probe_pairs(gpu_backend.dev(Batch(types_p, [pk, pval])))
    assert len(proc.probe_pairs(gpu_backend.dev(empty))) == 0
    fused = proc.fused_results()
    proc.close()
    hit = pk < 5000
    assert len(pairs) == int(hit.sum())
    assert fused[0]["value"] == int(hit.sum())
    assert fused[1]["value"] == int(sum(int(v) for v in pval[hit]))
    assert fused[2]["value"] == int(pval[hit].min())


def test_join_multi_batch_build_and_fused(gpu_backend):
    """Build from three batches (ordinals continue across them), duplicate
    keys (chains in build order), fused COUNT(*) / SUM / MIN / MAX."""
    rng = np.random.default_rng(11)
    types_b = [T(SIGNED, 8), T(SIGNED, 8)]
    types_p = [T(SIGNED, 8), T(SIGNED, 8, nullable=True)]
    batches = []
    for i in range(3):
        k = rng.integers(0, 3000, 4000).astype(np.int64)
        w = rng.integers(-100, 100, 4000).astype(np.int64)
        batches.append(Batch(types_b, [k, w]))
    pk = rng.integers(0, 3500, 100000).astype(np.int64)
    pval = rng.integers(-10 ** 15, 10 ** 15, 100000).astype(np.int64)
    pvv = rng.random(100000) > 0.1
    probe = Batch(types_p, [pk, pval], [None, pvv])
    case = JoinCase(types_b, types_p, [0], [0], fused=[(AGG_COUNT_STAR, -1), (AGG_SUM, 1), (AGG_MIN, 1), (AGG_MAX, -3),
                                                        ])
    kern, proc = gpu_backend.join_processor(case)
    for b in batches:
        proc.append(gpu_backend.dev(b))
    proc.build()
    pairs = proc.probe_pairs(gpu_backend.dev(probe))
    fused = proc.fused_results()
    oj = O.OracleJoin(case.to_c())
    for b in batches:
        oj.append(OracleBackend().host(b))
    oj.build()
    opairs = oj.probe(OracleBackend().host(probe))
    srt = lambda p: p[np.lexsort((p[:, 1], p[:, 0]))]
    assert np.array_equal(srt(pairs), srt(opairs))
    rows = (opairs[:, 0] + (1 << 32)).astype(np.int64)
    bw = np.concatenate([b.columns[1] for b in batches])
    vals = [int(pval[r]) for r in rows if pvv[r]]
    assert fused[0]["value"] == len(opairs)
    assert fused[1]["value"] == sum(vals) and fused[1]["count"] == len(vals)
    assert fused[2]["value"] == min(vals)
    assert fused[3]["value"] == int(bw[opairs[:, 1]].max())
    proc.close()


def test_filter_and_project_random(gpu_backend):
    rng = np.random.default_rng(3)
    n = 100003
    types = [T(SIGNED, 8, nullable=True), T(DECIMAL, 8, 2, True, 15), T(FLOAT, 8, nullable=True), T(SIGNED, 2)]
    cols = [rng.integers(-50, 50, n).astype(np.int64), rng.integers(-10 ** 6, 10 ** 6, n).astype(np.int64),
            rng.normal(0, 10, n), rng.integers(-100, 100, n).astype(np.int16)]
    valid = [rng.random(n) > 0.1, rng.random(n) > 0.1, rng.random(n) > 0.1, None]
    sel = (rng.random(n) > 0.3).astype(np.uint8)
    E = cases._abi()
    p = cases.Prog(types)
    a = p.cmp(E.CMP_GT, p.col(0), p.const(3, T(SIGNED, 8)))
    b = p.cmp(E.CMP_LE, p.col(2), p.const_f64(1.5))
    r = p._emit(E.EXPR_AND, E.REG_BOOL, T(3, 1, nullable=True), a=a, b=b)
    batch = Batch(types, cols, valid, sel)
    assert np.array_equal(gpu_backend.filter(types, p, r, batch), OracleBackend().filter(types, p, r, batch))
    # project: i128 decimal products, f64 arithmetic, a bool output
    q1 = cases.Prog(types)
    x = q1.op(E.EXPR_MUL, q1.col(1), q1.op(E.EXPR_SUB, q1.const(100, T(DECIMAL, 8, 2, False, 3)), q1.col(1), T(DECIMAL, 8, 2, True, 16)), T(DECIMAL, 16, 4, True, 31))
    q2 = cases.Prog(types)
    y = q2.op(E.EXPR_ADD, q2.col(2), q2.const_f64(0.25), T(FLOAT, 8, nullable=True), cls=E.REG_F64)
    q3 = cases.Prog(types)
    z = q3.cmp(E.CMP_NE, q3.col(3), q3.const(0, T(SIGNED, 8)))
    outs = [(q1, x, T(DECIMAL, 16, 4, True, 31)), (q2, y, T(FLOAT, 8, nullable=True)), (q3, z, T(3, 1))]
    g = gpu_backend.project(types, outs, batch)
    o = OracleBackend().project(types, outs, batch)
    for (gv, gm), (ov, om) in zip(g, o):
        mask = sel.astype(bool)
        gv, ov = np.asarray(gv, dtype=object), np.asarray(ov, dtype=object)
        assert np.array_equal(gv[mask], ov[mask])
        if om is not None:
            assert np.array_equal(gm[mask], om[mask])


def test_project_store_forms_generated_equals_interpreter(gpu_backend, monkeypatch):
    """Every output store form of a projection (i128 decimal, f64, f32, a
    BOOL bitmap, 2- and 8-byte integers, validity bitmaps) under an input
    selection, over a ragged row count: the kernel generated for the
    projection equals the interpreter kernel (QDB_GPU_JIT=0) and the oracle
    on every selected row."""
    rng = np.random.default_rng(11)
    n = 70001
    types = [T(SIGNED, 8, nullable=True), T(DECIMAL, 8, 2, True, 15), T(FLOAT, 8, nullable=True), T(SIGNED, 2)]
    cols = [rng.integers(-50, 50, n).astype(np.int64), rng.integers(-10 ** 6, 10 ** 6, n).astype(np.int64),
            rng.normal(0, 10, n), rng.integers(-100, 100, n).astype(np.int16)]
    valid = [rng.random(n) > 0.1, rng.random(n) > 0.1, rng.random(n) > 0.1, None]
    sel = (rng.random(n) > 0.3).astype(np.uint8)
    batch = Batch(types, cols, valid, sel)
    E = cases._abi()

    def outputs():
        q1 = cases.Prog(types)
        x = q1.op(E.EXPR_MUL, q1.col(1), q1.const(100, T(DECIMAL, 8, 2, False, 3)), T(DECIMAL, 16, 4, True, 31))
        q2 = cases.Prog(types)
        y = q2.op(E.EXPR_ADD, q2.col(2), q2.const_f64(0.25), T(FLOAT, 8, nullable=True), cls=E.REG_F64)
        q3 = cases.Prog(types)
        y4 = q3.op(E.EXPR_MUL, q3.col(2), q3.const_f64(3.0), T(FLOAT, 4, nullable=True), cls=E.REG_F64)
        q4 = cases.Prog(types)
        z = q4.cmp(E.CMP_NE, q4.col(3), q4.const(0, T(SIGNED, 8)))
        q5 = cases.Prog(types)
        w = q5.col(3)
        q6 = cases.Prog(types)
        v = q6.op(E.EXPR_ADD, q6.col(0), q6.const(7, T(SIGNED, 8)), T(SIGNED, 8, nullable=True))
        return [(q1, x, T(DECIMAL, 16, 4, True, 31)), (q2, y, T(FLOAT, 8, nullable=True)), (q3, y4, T(FLOAT, 4, nullable=True)),
                (q4, z, T(3, 1)), (q5, w, T(SIGNED, 2)), (q6, v, T(SIGNED, 8, nullable=True))]

    gen = gpu_backend.project(types, outputs(), batch)
    monkeypatch.setenv("QDB_GPU_JIT", "0")
    interp = gpu_backend.project(types, outputs(), batch)
    monkeypatch.delenv("QDB_GPU_JIT")
    orc = OracleBackend().project(types, outputs(), batch)
    mask = sel.astype(bool)
    for k, ((gv, gm), (iv, im), (ov, om)) in enumerate(zip(gen, interp, orc)):
        gv, iv, ov = (np.asarray(a, dtype=object)[mask] for a in (gv, iv, ov))
        assert np.array_equal(gv, iv), f"output {k}: generated != interpreter"
        assert np.array_equal(gv, ov), f"output {k}: generated != oracle"
        if om is not None:
            assert np.array_equal(gm[mask], im[mask]) and np.array_equal(gm[mask], om[mask]), f"output {k}: validity"


def test_group_hash_matches_rh_hash(gpu_backend):
    """The group-hash entry point computes the reference's rh_hash."""
    q = gpu_backend.q
    rng = np.random.default_rng(1)
    n = 5000
    types = [T(SIGNED, 8), T(SIGNED, 4, nullable=True), T(FLOAT, 8)]
    cols = [rng.integers(-10 ** 12, 10 ** 12, n).astype(np.int64), rng.integers(-5, 5, n).astype(np.int32), rng.normal(0, 1, n)]
    cols[2][:10] = [0.0, -0.0, np.nan, -np.nan, 1.0, -1.0, np.inf, 2.0, 3.0, 4.0]
    valid = [None, rng.random(n) > 0.2, None]
    batch = Batch(types, cols, valid)
    d = gpu_backend.dev(batch)
    hb = OracleBackend().host(batch)
    for keys in ([0], [1, 0, 2], [2]):
        fn = C.c_void_p()
        slot = q.check(q.lib().qdb_gpu_compile_group_hash(C.byref(gpu_backend.cast(cases.schema_c(types))), len(keys), C.byref(fn)), "hash")
        f = q.abi.GROUP_HASH_FN(fn.value)
        out = gpu_backend.H.DeviceBuffer(n * 8)
        kc = (C.c_int64 * len(keys))(*keys)
        assert f(C.byref(d.rs), out.ptr, kc)
        got = out.to_numpy(np.uint64, n)
        kk = (C.c_int32 * len(keys))(*keys)
        tt = (O.abi.QdbValueType * len(types))(*[cases.value_type_c(t) for t in types])
        want = np.array([O.LIB.oracle_key_hash(hb.rs.Columns, tt, kk, len(keys), i) for i in range(n)], dtype=np.uint64)
        assert np.array_equal(got, want), keys
        q.lib().qdb_gpu_release_slot(q.abi.SLOT_HASH, slot)


def test_stage_rowset_round_trip_and_aggregate(gpu_backend):
    """qdb_gpu_stage_rowset copies a host batch (8-, 4-, 1-byte and 16-byte
    columns, a bit-packed bool, validity masks, a selection and hashes) into
    HBM through pinned memory; the staged batch downloads to the same bytes
    and aggregates to the oracle's result on the host batch."""
    q = gpu_backend.q
    rng = np.random.default_rng(3)
    n = 100_003
    types = [T(SIGNED, 8), T(SIGNED, 4, nullable=True), T(SIGNED, 1), T(DECIMAL, 16, 2, True, 30), T(3, 1)]
    cols = [rng.integers(-10 ** 12, 10 ** 12, n).astype(np.int64), rng.integers(-50, 50, n).astype(np.int32),
            rng.integers(-3, 4, n).astype(np.int8), [int(x) * 10 ** 20 + 7 for x in rng.integers(-10 ** 6, 10 ** 6, n)],
            rng.random(n) > 0.5]
    valid = [None, rng.random(n) > 0.2, None, rng.random(n) > 0.3, None]
    sel = (rng.random(n) > 0.25).astype(np.uint8)
    hashes = rng.integers(0, 1 << 62, n).astype(np.uint64)
    hb = O.HostRowSet(types, cols, valid, selection=sel, hashes=hashes)
    schema = gpu_backend.cast(cases.schema_c(types))
    rs = q.abi.TRowSet.from_buffer_copy(hb.rs)
    staged = q.abi.TRowSet()
    q.check(q.lib().qdb_gpu_stage_rowset(C.byref(rs), C.byref(schema), C.byref(staged)), "stage")
    assert staged.RowCount == n and staged.Selection and staged.Hash and staged.Destroy
    vts = [gpu_backend.H.ValueType(t.kind, t.width, t.scale, t.nullable) for t in types]
    back = gpu_backend.H.download_rowset(staged, vts)
    want = O.read_host_rowset(hb.rs, types)
    for c, ((gv, gm), (wv, wm)) in enumerate(zip(back, want)):
        assert list(gv) == list(wv), f"column {c}"
        if wm is not None:
            assert np.array_equal(gm, wm), f"mask {c}"
    got_sel = np.empty(n, np.uint8)
    q.check(q.lib().qdb_gpu_memcpy_d2h(got_sel.ctypes.data, staged.Selection, n), "sel")
    got_hash = np.empty(n, np.uint64)
    q.check(q.lib().qdb_gpu_memcpy_d2h(got_hash.ctypes.data, staged.Hash, n * 8), "hash")
    assert np.array_equal(got_sel, sel) and np.array_equal(got_hash, hashes)
    # The staged batch is a kernel input like any device batch.
    case = AggCase(types, [2], [(AGG_COUNT_STAR, -1), (AGG_SUM, 0), (AGG_SUM, 3), (AGG_MIN, 1)])
    kern = q.abi.QdbAggregateKernels()
    q.check(q.lib().qdb_gpu_compile_aggregate(C.byref(gpu_backend.cast(case.to_c())), C.byref(kern)), "compile")
    p = q.lib().qdb_gpu_agg_processor_new(C.byref(kern), 8, 0, 0)
    assert q.lib().qdb_gpu_agg_processor_add(p, C.byref(staged)) == 0
    out = q.abi.TRowSet()
    assert q.lib().qdb_gpu_agg_processor_finish(p, C.byref(out)) >= 0
    ot = [gpu_backend.H.ValueType(t.Kind, t.Width, t.Scale, bool(t.Nullable)) for t in kern.OutputTypes[: kern.NumOutputColumns]]
    got = sorted(result_rows(gpu_backend.H.download_rowset(out, ot)), key=sort_key)
    q.lib().qdb_gpu_rowset_release(C.byref(out))
    q.lib().qdb_gpu_agg_processor_free(p)
    q.lib().qdb_gpu_release_slot(q.abi.SLOT_AGGREGATE, kern.Slot)
    q.lib().qdb_gpu_rowset_release(C.byref(staged))
    wantagg = sorted(result_rows(OracleBackend().aggregate(case, [Batch(types, cols, valid, sel)])), key=sort_key)
    assert got == wantagg


@pytest.mark.parametrize("nkeys,max_keys", [(3000, 1 << 14), (60000, 1 << 14), (5000, 100)])
def test_runtime_filter_equals_oracle(gpu_backend, nkeys, max_keys):
    """The runtime join filter published and probed on the GPU keeps exactly
    the rows the oracle's restatement of the reference keeps (exact set
    below MaxKeys distinct hashes, blocked Bloom filter above), with build
    and probe selections, repeated build hashes and a hash 0; its statistics
    (kind, exact keys, Bloom bits set) agree."""
    H = gpu_backend.H
    rng = np.random.default_rng(nkeys)
    build = rng.integers(0, 1 << 63, nkeys, dtype=np.uint64)
    build[:3] = [0, 1, 1]
    bsel = (rng.random(nkeys) > 0.2).astype(np.uint8)
    probe = np.concatenate([build[rng.integers(0, nkeys, 50000)], rng.integers(0, 1 << 63, 150000, dtype=np.uint64)])
    probe[:2] = [0, 2]
    psel = (rng.random(len(probe)) > 0.1).astype(np.uint8)
    f = H.RuntimeFilter(max_keys=max_keys)
    half = nkeys // 2
    for a, b in ((0, half), (half, nkeys)):
        f.build(H.DeviceBuffer.from_numpy(build[a:b]), b - a, H.DeviceBuffer.from_numpy(bsel[a:b]))
    kind = f.publish()
    of = O.OracleRuntimeFilter(max_keys=max_keys)
    of.build(build, bsel)
    assert kind == of.publish() == (1 if len(np.unique(build[bsel.astype(bool)])) <= max_keys else 2)
    out, kept = f.probe(H.DeviceBuffer.from_numpy(probe), len(probe), H.DeviceBuffer.from_numpy(psel))
    got = out.to_numpy(np.uint8, len(probe))
    want, wkept = of.probe(probe, psel)
    bad = np.nonzero(got != want)[0]
    assert len(bad) == 0 and kept == wkept, (
        f"{len(bad)} rows differ (first {bad[:5].tolist()}: got {got[bad[:5]].tolist()}, want {want[bad[:5]].tolist()}, "
        f"hashes {[hex(int(x)) for x in probe[bad[:5]]]}), kept {kept} vs {wkept}, stats {f.stats()}")
    st = f.stats()
    if kind == 1:
        assert st["exact_keys"] == len(np.unique(build[bsel.astype(bool)]))
    else:
        assert st["bloom_words"] == (2 << 20) // 8 and st["bloom_bits"] > 0
    f.close()


def test_runtime_filter_on_the_join_keeps_every_match(gpu_backend):
    """The filtered-build join with its runtime filter: build hashes are the
    group-hash entry point's rh_hash of the selected build keys (segment = 1,
    60K of 300K rows: a Bloom filter), the probe rows the filter rejects are
    deselected before the probe, and the join's pairs and fused COUNT/SUM
    equal those without the filter; the filter rejects most non-matching
    probe rows."""
    from qumirdb_amd import workloads as W
    from qumirdb_amd.host import FilterKernel, GroupHash, RowJoinKernels, TRowJoinProcessor, RuntimeFilter
    B, Pn = 300_000, 2_000_000
    jd = W.JoinDevice(B, Pn)
    bsel = FilterKernel(W.BUILD_TYPES, *W.join_build_predicate())(jd.build)
    jd.build.set_selection_buffer(bsel)
    bh = GroupHash(W.BUILD_TYPES, 1)(jd.build, [W.BUILD_KEY])
    ph = GroupHash(W.PROBE_TYPES, 1)(jd.probe, [W.PROBE_FK])
    rf = RuntimeFilter()
    rf.build(bh, B, bsel)
    assert rf.publish() == 2
    keep, kept = rf.probe(ph, Pn)

    def run(sel):
        k = RowJoinKernels(W.join_spec(fused=True))
        p = TRowJoinProcessor(k, build_is_left=False)
        p.append(jd.build)
        p.build()
        probe = jd.probe
        if sel is not None:
            probe = jd.probe.view(0, Pn)
            probe.set_selection_buffer(sel)
        pairs = p.probe_pairs(probe)
        fused = p.fused_results()
        p.close()
        return pairs[np.lexsort((pairs[:, 1], pairs[:, 0]))], fused

    plain, pf = run(None)
    filtered, ff = run(keep)
    assert np.array_equal(plain, filtered) and pf == ff
    assert len(plain) <= kept < len(plain) + (Pn - len(plain)) // 10


def test_group_hash_known_answers(gpu_backend):
    """The group-hash entry point returns the reference's rh_hash values,
    worked out by hand from qdb/kernel/gen.cpp:233-379 over the key struct
    (a NOT NULL BIGINT 1; a nullable BIGINT 5 and NULL)."""
    q = gpu_backend.q
    for types, cols, valid, want in (
            ([T(SIGNED, 8)], [np.array([1], np.int64)], None, [0xE61C480508B75932]),
            ([T(SIGNED, 8, nullable=True)], [np.array([5, 5], np.int64)], [np.array([True, False])],
             [0x33D914A28E354A67, 0xCD94BF3ECEF65C4F])):
        d = gpu_backend.dev(Batch(types, cols, valid))
        fn = C.c_void_p()
        slot = q.check(q.lib().qdb_gpu_compile_group_hash(C.byref(gpu_backend.cast(cases.schema_c(types))), 1, C.byref(fn)), "hash")
        out = gpu_backend.H.DeviceBuffer(len(want) * 8)
        assert q.abi.GROUP_HASH_FN(fn.value)(C.byref(d.rs), out.ptr, (C.c_int64 * 1)(0))
        assert [int(x) for x in out.to_numpy(np.uint64, len(want))] == want
        q.lib().qdb_gpu_release_slot(q.abi.SLOT_HASH, slot)


def test_group_by_float_keys_zero_and_nan(gpu_backend):
    """GROUP BY a DOUBLE key holding +0, -0 and NaNs of several bit patterns:
    +0/-0 form one group and every NaN one group (canonical float key bits,
    qdb/kernel/gen.cpp:192-231), as in the oracle; keys are compared by
    their canonical bits, since the group's stored key is whichever
    representative arrived first."""
    rng = np.random.default_rng(7)
    n = 20_000
    nan2 = np.frombuffer(np.uint64(0x7FF0000000000123).tobytes(), np.float64)[0]
    nan3 = np.frombuffer(np.uint64(0xFFF8000000000000).tobytes(), np.float64)[0]
    pool = np.array([0.0, -0.0, np.nan, nan2, nan3, 1.5, -2.25, np.inf, -np.inf])
    key = pool[rng.integers(0, len(pool), n)]
    v = rng.integers(-1000, 1000, n).astype(np.int64)
    types = [T(FLOAT, 8), T(SIGNED, 8)]
    case = AggCase(types, [0], [(AGG_COUNT_STAR, -1), (AGG_SUM, 1)])
    b = Batch(types, [key, v])

    def canon(rows):
        out = []
        for k, c, s in rows:
            k = 0.0 if k == 0 else k
            out.append(("nan" if k != k else k, c, s))
        return sorted(out, key=lambda r: (str(r[0]), r[1]))

    got = canon(result_rows(gpu_backend.aggregate(case, [b])))
    want = canon(result_rows(OracleBackend().aggregate(case, [b])))
    assert got == want
    assert len(got) == 6


# ---------------------------------------------------------------------------
# The three workloads
# ---------------------------------------------------------------------------

def lineitem_oracle(rows):
    from qumirdb_amd import workloads as W
    from qumirdb_amd.host import schema_c
    abi = O.abi
    cols = O.gen_lineitem(W.SEED_LINEITEM, 0, rows)
    hrs = O.HostRowSet(W.LINEITEM_TYPES, cols)
    fs = abi.QdbFilterSpec()
    fs.Input = schema_c(W.LINEITEM_TYPES)
    p, r = W.lineitem_predicate()
    fs.Predicate = p.build(r)
    sel = O.filter_rows(fs, hrs)
    ps = abi.QdbProjectSpec()
    ps.Input = schema_c(W.LINEITEM_TYPES)
    outs = [W.lineitem_disc_price(), W.lineitem_charge()]
    ps.NumOutputs = 2
    for k, (pp, rr) in enumerate(outs):
        ps.OutputTypes[k] = pp.result_type(rr).to_c()
        ps.Outputs[k] = pp.build(rr)
    proj = O.project_rows(ps, hrs, [pp.result_type(rr) for pp, rr in outs])
    full = O.HostRowSet(W.lineitem_projected_types(), cols + [proj[0][0], proj[1][0]], selection=sel)
    return W.lineitem_finalize(O.aggregate(W.lineitem_plain_spec().to_c(), [full]))


@pytest.mark.parametrize("rows", [1000, 6_000_000])
def test_lineitem_fused_vs_oracle(gpu_backend, rows):
    """WHERE shipdate <= 2436 GROUP BY flag, status with the four exact
    decimal sums, two AVGs and COUNT(*): the fused kernel equals the
    oracle's filter -> project -> aggregate bit for bit (AVG as f64)."""
    from qumirdb_amd import workloads as W
    from qumirdb_amd.host import AggregateKernels, TAggregateProcessor
    dev = W.LineitemDevice(rows)
    ak = AggregateKernels(W.lineitem_fused_spec())
    proc = TAggregateProcessor(ak, initial_capacity=8)
    proc.add(dev.rowset)
    got = W.lineitem_finalize_on_device(proc, ak)
    assert got == lineitem_oracle(rows)
    assert len(got) == 6


def test_lineitem_operators_equal_fused(gpu_backend):
    """The drop-in operator chain (filter kernel -> project kernel ->
    aggregate) equals the fused kernel."""
    from qumirdb_amd import workloads as W
    from qumirdb_amd.host import (AggregateKernels, DeviceRowSet, FilterKernel, ProjectKernel, TAggregateProcessor)
    rows = 500_000
    dev = W.LineitemDevice(rows)
    sel = FilterKernel(W.LINEITEM_TYPES, *W.lineitem_predicate())(dev.rowset)
    proj = ProjectKernel(W.LINEITEM_TYPES, [W.lineitem_disc_price(), W.lineitem_charge()])
    bufs, _ = proj(dev.rowset)
    full = DeviceRowSet(W.lineitem_projected_types(), dev.bufs + bufs[:2], rows=rows)
    full.set_selection_buffer(sel)
    k1 = AggregateKernels(W.lineitem_plain_spec())
    p1 = TAggregateProcessor(k1, initial_capacity=8)
    p1.add(full)
    plain = W.lineitem_finalize_on_device(p1, k1)
    k2 = AggregateKernels(W.lineitem_fused_spec())
    p2 = TAggregateProcessor(k2, initial_capacity=8)
    p2.add(dev.rowset)
    assert plain == W.lineitem_finalize_on_device(p2, k2)


def run_join(build_rows, probe_rows, table_mode=0):
    from qumirdb_amd import workloads as W
    from qumirdb_amd.host import FilterKernel, RowJoinKernels, TRowJoinProcessor
    jd = W.JoinDevice(build_rows, probe_rows)
    sel = FilterKernel(W.BUILD_TYPES, *W.join_build_predicate())(jd.build)
    jd.build.set_selection_buffer(sel)
    proc = TRowJoinProcessor(RowJoinKernels(W.join_spec(fused=True, table_mode=table_mode)), build_is_left=False)
    proc.append(jd.build)
    proc.build()
    pairs = proc.probe_pairs(jd.probe)
    return jd, sel, pairs, proc.fused_results(), proc


@pytest.mark.parametrize("table_mode", [0, 1])
def test_join_small_vs_oracle(gpu_backend, table_mode):
    """The join workload at 150K x 2M against the oracle's row-table join:
    the same sorted pair set, fused COUNT(*) and exact SUM; with the table
    the build picks (dense direct table here) and with the hash table."""
    from qumirdb_amd import workloads as W
    B, P = 150_000, 2_000_000
    jd, sel, pairs, fused, proc = run_join(B, P, table_mode)
    key, seg = O.gen_join_build(W.SEED_JOIN, 0, B, B)
    fk, val = O.gen_join_probe(W.SEED_JOIN, 0, P, B)
    oj = O.OracleJoin(W.join_spec().to_c())
    oj.append(O.HostRowSet(W.BUILD_TYPES, [key, seg], selection=(seg == 1)))
    oj.build()
    op = oj.probe(O.HostRowSet(W.PROBE_TYPES, [fk, val]))
    srt = lambda p: p[np.lexsort((p[:, 1], p[:, 0]))]
    assert np.array_equal(srt(pairs), srt(op))
    rows = (op[:, 0] + (1 << 32)).astype(np.int64)
    assert fused[0]["value"] == len(op)
    assert fused[1]["value"] == int(val[rows].sum(dtype=np.int64))
    assert abs(len(op) / P - 0.2) < 0.01


@pytest.mark.parametrize("table_mode", [0, 1])
def test_join_full_size_properties(gpu_backend, table_mode):
    """The 1.5M x 60M join: every pair joins equal keys of a segment-1 build
    row; the pair count, the fused COUNT(*) and the rows whose key is in
    the filtered build agree; the fused exact SUM equals an independent
    aggregate over the probe rows the pairs name."""
    from qumirdb_amd import workloads as W
    B, P = 1_500_000, 60_000_000
    jd, sel, pairs, fused, proc = run_join(B, P, table_mode)
    key = jd.key.to_numpy(np.int64, B)
    seg = jd.segment.to_numpy(np.int32, B)
    fk = jd.fk.to_numpy(np.int64, P)
    val = jd.value.to_numpy(np.int64, P)
    rows = (pairs[:, 0] + (1 << 32)).astype(np.int64)
    assert len(np.unique(rows)) == len(rows)          # build keys are unique
    assert np.array_equal(fk[rows], key[pairs[:, 1]])
    assert (seg[pairs[:, 1]] == 1).all()
    in_build = np.zeros(B + 1, bool)
    in_build[key[seg == 1]] = True
    assert len(rows) == int(in_build[fk].sum())
    assert fused[0]["value"] == len(rows)
    assert fused[1]["value"] == sum(int(x) for x in np.add.reduceat(val[np.sort(rows)], np.arange(0, len(rows), 1 << 20)))


@pytest.mark.parametrize("rows,distinct", [(1_000_000, 100_000), (20_000_000, 10_000_000)])
def test_groups_vs_oracle(gpu_backend, rows, distinct):
    """High-cardinality Zipf group-by: keys, counts, i64 sums, MIN/MAX bit
    exact; the f64 SUM equals the oracle's correctly rounded exact sum and
    is within 1e-12 relative of its Kahan-Babuska-Neumaier sum."""
    from qumirdb_amd import workloads as W
    from qumirdb_amd.host import AggregateKernels, TAggregateProcessor
    cdf = W.zipf1_cdf(distinct)
    dev = W.GroupsDevice(rows, distinct, cdf=cdf)
    proc = TAggregateProcessor(AggregateKernels(W.groups_spec()), initial_capacity=1 << 20)
    proc.add(dev.rowset)
    got = proc.finish()
    key, vi, vf = O.gen_groups(W.SEED_GROUPS, 0, rows, cdf)
    hb = O.HostRowSet(W.GROUPS_TYPES, [key, vi, vf])
    exact = O.aggregate(W.groups_spec().to_c(), [hb], fsum_mode=2)
    kahan = O.aggregate(W.groups_spec().to_c(), [hb], fsum_mode=1)
    gi, ei, ki = np.argsort(got[0][0]), np.argsort(exact[0][0]), np.argsort(kahan[0][0])
    for c in range(5):
        assert np.array_equal(np.asarray(got[c][0])[gi], np.asarray(exact[c][0])[ei]), c
    gs = np.asarray(got[5][0])[gi]
    ks = np.asarray(kahan[5][0])[ki]
    rel = np.abs(gs - ks) / np.maximum(np.abs(ks), 1e-300)
    assert (rel <= F64_REL_TOL).all(), float(rel.max())
    assert np.array_equal(gs, np.asarray(exact[5][0])[ei])


def test_groups_full_size_sampled_vs_oracle(gpu_backend):
    """The high-cardinality group-by at its full size (100M rows, Zipf(1)
    keys over 10M values, 7.2M groups; the bucketed path) against the oracle:
    every group whose key falls in a 1/32 sample (by a multiplicative hash of
    the key) is compared bit for bit (key, COUNT, SUM(vi), MIN/MAX(vf), and
    the f64 SUM against the correctly rounded exact sum, within 1e-12 of
    Kahan), the oracle aggregating exactly the sampled keys' rows; and over
    all groups the counts add up to the row count and the i64 sums to the
    column's total."""
    from qumirdb_amd import workloads as W
    from qumirdb_amd.host import AggregateKernels, TAggregateProcessor
    rows, distinct = 100_000_000, 10_000_000
    cdf = W.zipf1_cdf(distinct)
    dev = W.GroupsDevice(rows, distinct, cdf=cdf)
    proc = TAggregateProcessor(AggregateKernels(W.groups_spec()), initial_capacity=1 << 24)
    proc.add(dev.rowset)
    got = proc.finish()
    del dev
    key, vi, vf = O.gen_groups(W.SEED_GROUPS, 0, rows, cdf)
    gk = np.asarray(got[0][0])
    assert int(np.asarray(got[1][0]).sum()) == rows
    assert int(np.asarray(got[2][0]).astype(np.int64).sum()) == int(vi.sum())
    assert len(gk) == len(np.unique(key))

    def sampled(k):
        return ((k.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)) >> np.uint64(59)) == 0

    rs = sampled(key)
    hb = O.HostRowSet(W.GROUPS_TYPES, [key[rs], vi[rs], vf[rs]])
    del key, vi, vf
    exact = O.aggregate(W.groups_spec().to_c(), [hb], fsum_mode=2)
    kahan = O.aggregate(W.groups_spec().to_c(), [hb], fsum_mode=1)
    gs = sampled(gk)
    gi = np.argsort(gk[gs])
    ei, ki = np.argsort(exact[0][0]), np.argsort(kahan[0][0])
    assert gs.sum() == len(exact[0][0]) > 100_000
    for c in range(6):
        g = np.asarray(got[c][0])[gs][gi]
        e = np.asarray(exact[c][0])[ei]
        assert np.array_equal(g.view(np.int64) if g.dtype == np.float64 else g,
                              e.view(np.int64) if e.dtype == np.float64 else e), c
    ks = np.asarray(kahan[5][0])[ki]
    g5 = np.asarray(got[5][0])[gs][gi]
    assert (np.abs(g5 - ks) <= F64_REL_TOL * np.maximum(np.abs(ks), 1e-300)).all()


def test_f64_sum_edge_values(gpu_backend):
    """f64 SUM beyond the usual range, per group:
    0: heavy cancellation inside the exact window, whose values have every
       bit in [2^-64, 2^64) (pairs of +-(2^60 + k) and small dyadic terms
       k * 2^-40): equal to the correctly rounded exact sum, and the same bits
       on a second run;
    1: values in [1e20, 1e25] (above the exact window), 2: values in
       [1e-25, 1e-22] (below it) and 6: N(0, 1e-6) values (bits below
       2^-64): within 1e-12 relative of Kahan;
    3: a +inf among finite values: +inf; 4: a NaN: NaN; 5: +inf and -inf: NaN.
    MIN/MAX of every group equal the oracle's (no NaN or signed zero in the
    groups that have a finite result)."""
    rng = np.random.default_rng(5)
    m = 20_000
    g0a = (rng.integers(0, 1 << 20, m).astype(np.float64) + 2.0 ** 60)
    g0 = np.concatenate([g0a, -g0a, rng.integers(-(1 << 30), 1 << 30, m).astype(np.float64) * 2.0 ** -40])
    g6 = rng.normal(0, 1e-6, m) + 1e-3
    g1 = 10.0 ** rng.uniform(20, 25, m)
    g2 = 10.0 ** rng.uniform(-25, -22, m)
    g3 = np.concatenate([rng.normal(0, 1, 1000), [np.inf]])
    g4 = np.concatenate([rng.normal(0, 1, 1000), [np.nan]])
    g5 = np.concatenate([rng.normal(0, 1, 1000), [np.inf, -np.inf]])
    groups = [g0, g1, g2, g3, g4, g5, g6]
    keys = np.concatenate([np.full(len(g), i, np.int64) for i, g in enumerate(groups)])
    vals = np.concatenate(groups)
    perm = rng.permutation(len(vals))
    keys, vals = keys[perm], vals[perm]
    types = [T(SIGNED, 8), T(FLOAT, 8)]
    case = AggCase(types, [0], [(AGG_SUM, 1), (AGG_MIN, 1), (AGG_MAX, 1), (AGG_COUNT_STAR, -1)])
    b = Batch(types, [keys, vals])
    got = sorted(result_rows(gpu_backend.aggregate(case, [b])), key=sort_key)
    again = sorted(result_rows(gpu_backend.aggregate(case, [b])), key=sort_key)
    exact = sorted(result_rows(OracleBackend(fsum_mode=2).aggregate(case, [b])), key=sort_key)
    kahan = sorted(result_rows(OracleBackend(fsum_mode=1).aggregate(case, [b])), key=sort_key)
    assert [r[0] for r in got] == list(range(7))
    assert got[0][1] == exact[0][1] and np.float64(got[0][1]).tobytes() == np.float64(again[0][1]).tobytes()
    for g in (1, 2, 6):
        assert abs(got[g][1] - kahan[g][1]) <= F64_REL_TOL * abs(kahan[g][1]), g
    assert got[3][1] == np.inf
    assert math.isnan(got[4][1]) and math.isnan(got[5][1])
    for g in (0, 1, 2, 3, 5, 6):
        assert got[g][2:] == exact[g][2:], g


def test_served_results_same_alone_and_in_the_mix(gpu_backend):
    """Requests served with continuous batching (admitted on arrival,
    dispatched in units to worker threads that each issue on their own
    stream; group-bys of a shape share one batched launch, joins of a tenant
    one replay of the captured probe pipeline) return exactly what each
    returns alone, one at a time."""
    from qumirdb_amd import serve as SV
    res = SV.Resident(lineitem_rows=200_000, build_rows=50_000, probe_rows=400_000, group_rows=300_000, distinct=40_000,
                      tenants=2)
    srv = SV.Server(res, full_results=True)
    try:
        stream = SV.make_stream(res, 48, 20_000.0, seed=7)
        mixed = srv.serve(SV.fresh(stream), batched=True)
        alone = srv.serve(SV.fresh(stream), batched=False)
        assert {r.shape for r in stream} == set(SV.SHAPES)
        assert {r.tenant for r in stream} == {0, 1}
        for a, b in zip(alone, mixed):
            assert a.result == b.result, f"request {a.rid} ({a.shape}, tenant {a.tenant}, rows {a.r0}..{a.r1})"
        # Requests ran together: some unit held more than one request, and
        # the units of different workers overlapped in time.
        spans = sorted((r.start, r.done) for r in mixed)
        assert any(s2 < d1 for (s1, d1), (s2, d2) in zip(spans, spans[1:]))
        # and a join request's fused aggregate against the oracle's pairs
        for j in [r for r in alone if r.shape == "join"][:2]:
            jd = res.tenants[j.tenant].jd
            fk = jd.fk.to_numpy(np.int64, jd.probe_rows)[j.r0:j.r1]
            key = jd.key.to_numpy(np.int64, jd.build_rows)
            seg = jd.segment.to_numpy(np.int32, jd.build_rows)
            val = jd.value.to_numpy(np.int64, jd.probe_rows)[j.r0:j.r1]
            build = set(key[seg == 1].tolist())
            hit = np.array([k in build for k in fk.tolist()], dtype=bool)
            assert j.result == (int(hit.sum()), int(hit.sum()), int(val[hit].sum()))
    finally:
        srv.close()
        res.close()


# ---------------------------------------------------------------------------
# Hash shuffle and partial aggregate
# ---------------------------------------------------------------------------

def _per_output(idx, offs):
    rows = idx.to_numpy(np.uint32, int(offs[-1])) if offs[-1] else np.zeros(0, np.uint32)
    return [rows[offs[o]:offs[o + 1]].tolist() for o in range(len(offs) - 1)]


def test_shuffle_reference_goldens(gpu_backend):
    """The reference's routing cases (test/test_scheduler_runtime_adapter.cpp:
    1496-1536, 1538-1562) through qdb_gpu_shuffle_partition, the batch's own
    Hash column giving the hashes."""
    H = gpu_backend.H

    def batch_with_hashes(h):
        d = gpu_backend.dev(Batch([T(SIGNED, 8)], [np.arange(len(h), dtype=np.int64)]))
        hb = H.DeviceBuffer.from_numpy(np.array(h, np.uint64))
        d.rs.Hash = hb.ptr
        d.buffers.append(hb)
        return d

    d = batch_with_hashes([0, 1, 2, 3, 5])
    assert _per_output(*H.shuffle_partition(d, 2)) == [[0, 2], [1, 3, 4]]
    assert _per_output(*H.shuffle_partition(d, 2, routes=[1, 0, 0, 1])) == [[1, 2, 4], [0, 3]]
    with pytest.raises(RuntimeError, match="partition 1, which it routes nowhere"):
        H.shuffle_partition(batch_with_hashes([0, 3]), 2, routes=[0, -1])
    # A routing table whose only bad entry no row reaches is fine.
    assert _per_output(*H.shuffle_partition(batch_with_hashes([0, 2]), 2, routes=[1, -1])) == [[], [0, 1]]


@pytest.mark.parametrize("outputs", [1, 3, 64, 300, 4096])
def test_shuffle_partition_matches_oracle(gpu_backend, outputs):
    """Routing of a 200K-row batch with a selection over a two-column key: the
    device rh_hash and the device partition equal the oracle's, row for row."""
    H = gpu_backend.H
    rng = np.random.default_rng(outputs)
    n = 200_003
    types = [T(SIGNED, 8), T(SIGNED, 4, nullable=True)]
    cols = [rng.integers(-10 ** 12, 10 ** 12, n).astype(np.int64), rng.integers(-1000, 1000, n).astype(np.int32)]
    valid = [None, rng.random(n) > 0.2]
    sel = (rng.random(n) > 0.3).astype(np.uint8)
    batch = Batch(types, cols, valid, sel)
    d = gpu_backend.dev(batch)
    vts = [H.ValueType(t.kind, t.width, t.scale, t.nullable) for t in types]
    gh = H.GroupHash(vts, 2)(d, [0, 1])
    hb = OracleBackend().host(batch)
    want_h = O.key_hashes(hb, [cases.value_type_c(t) for t in types], [0, 1])
    # The entry point hashes the selected rows only.
    keep = sel.astype(bool)
    assert np.array_equal(gh.to_numpy(np.uint64, n)[keep], want_h[keep])
    idx, offs = H.shuffle_partition(d, outputs, hashes=gh)
    want_rows, want_offs = O.shuffle_partition(want_h, sel, outputs)
    assert np.array_equal(offs, want_offs)
    assert np.array_equal(idx.to_numpy(np.uint32, int(offs[-1])), want_rows)


def test_shuffle_full_size_properties(gpu_backend):
    """20M rows over 4096 outputs (a 13-bit sort key): each output's count is
    the histogram of the selected rows' residues, each output's rows ascend,
    and together they are exactly the selected rows."""
    H = gpu_backend.H
    rng = np.random.default_rng(11)
    n, outputs = 20_000_000, 4096
    h = rng.integers(0, np.iinfo(np.uint64).max, n, dtype=np.uint64, endpoint=True)
    sel = (rng.random(n) > 0.25).astype(np.uint8)
    d = gpu_backend.dev(Batch([T(SIGNED, 8)], [np.zeros(n, np.int64)], None, sel))
    hb = H.DeviceBuffer.from_numpy(h)
    idx, offs = H.shuffle_partition(d, outputs, hashes=hb)
    res = (h % np.uint64(outputs)).astype(np.int64)
    counts = np.bincount(res[sel.astype(bool)], minlength=outputs)
    assert np.array_equal(np.diff(offs), counts)
    rows = idx.to_numpy(np.uint32, int(offs[-1])).astype(np.int64)
    assert np.array_equal(np.sort(rows), np.flatnonzero(sel))
    out_of = res[rows]
    assert np.all(np.diff(out_of) >= 0)
    same_out = np.diff(out_of) == 0
    assert np.all(np.diff(rows)[same_out] > 0)


def test_gather_rowset_equals_host_take(gpu_backend):
    """Gathering each shuffle output of a batch view (bit offsets on the bool
    column and the validity bitmaps) gives the host's rows at those indexes:
    every column, the validity, the hashes and the selection."""
    H = gpu_backend.H
    rng = np.random.default_rng(21)
    n, r0, outputs = 50_000, 37, 7
    types = [T(SIGNED, 8), T(SIGNED, 4, nullable=True), T(DECIMAL, 16, 2, True, 30), T(3, 1), T(SIGNED, 1), T(FLOAT, 8)]
    cols = [rng.integers(-10 ** 12, 10 ** 12, n).astype(np.int64), rng.integers(-50, 50, n).astype(np.int32),
            [int(x) * 10 ** 20 + 7 for x in rng.integers(-10 ** 6, 10 ** 6, n)], rng.random(n) > 0.5,
            rng.integers(-3, 4, n).astype(np.int8), rng.normal(0, 1, n)]
    valid = [None, rng.random(n) > 0.2, rng.random(n) > 0.3, None, None, None]
    sel = (rng.random(n) > 0.4).astype(np.uint8)
    full = gpu_backend.dev(Batch(types, cols, valid, sel))
    d = full.view(r0, n)
    h = rng.integers(0, 1 << 62, n - r0).astype(np.uint64)
    hb = H.DeviceBuffer.from_numpy(h)
    d.rs.Hash = hb.ptr
    idx, offs = H.shuffle_partition(d, outputs)
    vts = [H.ValueType(t.kind, t.width, t.scale, t.nullable) for t in types]
    host = OracleBackend().host(Batch(types, cols, valid, sel))  # holds the host buffers
    want_all = O.read_host_rowset(host.rs, types)
    total = 0
    for o in range(outputs):
        cnt = int(offs[o + 1] - offs[o])
        rows = idx.to_numpy(np.uint32, int(offs[-1]))[offs[o]:offs[o + 1]].astype(np.int64)
        g = H.gather_rowset(d, idx, int(offs[o]), cnt)
        assert g.rows == cnt
        got = H.download_rowset(g.rs, vts)
        for c, ((gv, gm), (wv, wm)) in enumerate(zip(got, want_all)):
            wv_rows = [wv[r0 + r] for r in rows]
            assert list(gv) == list(wv_rows), f"output {o} column {c}"
            if wm is not None:
                assert np.array_equal(gm, np.asarray(wm)[r0 + rows]), f"output {o} mask {c}"
        if cnt:
            gs = np.empty(cnt, np.uint8)
            gpu_backend.q.check(gpu_backend.q.lib().qdb_gpu_memcpy_d2h(gs.ctypes.data, g.rs.Selection, cnt), "sel")
            gh = np.empty(cnt, np.uint64)
            gpu_backend.q.check(gpu_backend.q.lib().qdb_gpu_memcpy_d2h(gh.ctypes.data, g.rs.Hash, cnt * 8), "hash")
            assert gs.all() and np.array_equal(gh, h[rows])
            assert np.all(h[rows] % np.uint64(outputs) == o)
        total += cnt
        g.close()
    assert total == int(sel[r0:].sum())


def _compile_agg(gpu_backend, case):
    H = gpu_backend.H
    kern = H.AggregateKernels.__new__(H.AggregateKernels)
    kern.spec = None
    kern.c_spec = gpu_backend.cast(case.to_c())
    kern.k = gpu_backend.q.abi.QdbAggregateKernels()
    gpu_backend.q.check(gpu_backend.q.lib().qdb_gpu_compile_aggregate(C.byref(kern.c_spec), C.byref(kern.k)), "compile")
    kern.out_types = [H.ValueType(t.Kind, t.Width, t.Scale, bool(t.Nullable)) for t in kern.k.OutputTypes[: kern.k.NumOutputColumns]]
    return kern


def _selected_rows(cols, sel):
    rows = result_rows(cols)
    return [r for r, s in zip(rows, sel) if s] if sel is not None else rows


@pytest.mark.parametrize("distinct,sample,max_groups,rows", [
    (40, 50_000, 1 << 20, 20_000),          # reduces well: one table, kept to the end
    (10 ** 9, 50_000, 1 << 20, 20_000),     # unique keys: emitted at 60K rows, then states
    (25_000, 10 ** 9, 15_000, 20_000),      # bounded table: emits at 15K groups
])
def test_partial_aggregate_equals_oracle(gpu_backend, distinct, sample, max_groups, rows):
    """The device partial aggregate makes the oracle's decisions after the
    same batches and emits the same outputs (tables as sets of groups, states
    row for row); the final aggregate (COUNT -> SUM) over its device outputs
    equals the one-phase aggregate."""
    import test_oracle as TO
    H = gpu_backend.H
    rng = np.random.default_rng(distinct % 1000 + 3)
    types, case = TO._partial_cases()
    batches = TO._partial_batches(rng, types, 6, rows, distinct)
    kern = _compile_agg(gpu_backend, case)
    pp = H.TPartialAggregateProcessor(kern, initial_capacity=8, sample_rows=sample, max_groups=max_groups)
    spec = case.to_c()
    pa = O.PartialAggregate(spec, sample_rows=sample, max_groups=max_groups)
    be = OracleBackend()
    devs = [gpu_backend.dev(b) for b in batches]
    outs = []
    for b, d in zip(batches, devs):
        seen = len(pa.outputs)
        pa.add(be.host(b))
        pp.add(d)
        got = pp.take()
        want = pa.outputs[seen:]
        assert len(got) == len(want) and pp.passing_through == pa.passing
        for g, (kind, w) in zip(got, want):
            gcols = H.download_rowset(g.rs, kern.out_types)
            if kind == "table":
                assert sorted(result_rows(gcols), key=sort_key) == sorted(result_rows(w), key=sort_key)
            else:
                gsel = np.empty(g.rows, np.uint8)
                gpu_backend.q.check(gpu_backend.q.lib().qdb_gpu_memcpy_d2h(gsel.ctypes.data, g.rs.Selection, g.rows), "sel")
                assert np.array_equal(gsel, b.selection)
                assert _selected_rows(gcols, gsel) == result_rows(O.partial_states(spec, w))
            outs.append(g)
    last = pp.finish()
    assert (last is None) == pa.passing
    if last is not None:
        want = O.aggregate(spec, pa.table) if pa.table else []
        assert sorted(result_rows(H.download_rowset(last.rs, kern.out_types)), key=sort_key) == \
            sorted(result_rows(want), key=sort_key)
        outs.append(last)
    ptypes = TO._partial_types(spec)
    fk = _compile_agg(gpu_backend, TO._combine_case(case, ptypes))
    fp = H.TAggregateProcessor(fk, initial_capacity=1 << 12)
    for o in outs:
        fp.add(o)
    final = sorted(result_rows(fp.finish()), key=sort_key)
    direct = sorted(result_rows(be.aggregate(case, batches)), key=sort_key)
    assert final == direct
    for o in outs:
        o.close()
    fp.close()
    pp.close()
    gpu_backend.q.lib().qdb_gpu_release_slot(gpu_backend.q.abi.SLOT_AGGREGATE, fk.k.Slot)
    gpu_backend.q.lib().qdb_gpu_release_slot(gpu_backend.q.abi.SLOT_AGGREGATE, kern.k.Slot)


def test_partial_aggregate_without_projection_keeps_hashing(gpu_backend):
    """An aggregate with fused stages has no partial-state projection: the
    entry point says so, and the partial aggregate keeps its bounded table
    and never passes rows through, whatever its reduction."""
    from qumirdb_amd import workloads as W
    H = gpu_backend.H
    ak = H.AggregateKernels(W.lineitem_fused_spec())
    dev = W.LineitemDevice(10_000)
    with pytest.raises(RuntimeError, match="fused stages|no projection"):
        H.partial_states(ak, dev.rowset)
    pp = H.TPartialAggregateProcessor(ak, sample_rows=1, min_reduction=1 << 40)
    pp.add(dev.rowset)
    assert not pp.passing_through and pp.take() == []
    last = pp.finish()
    ref = H.TAggregateProcessor(ak)
    ref.add(dev.rowset)
    assert last is not None and last.rows == ref.finish_in_hbm() > 0
    ref.close()
    last.close()
    pp.close()
    ak.release()


@pytest.mark.parametrize("hot", ["1", "0"], ids=["hot_keys", "no_hot_keys"])
def test_bucketed_one_word_key_hot_keys_vs_oracle(gpu_backend, monkeypatch, hot):
    """The bucketed path with a one-word key whose rows are skewed (40% on
    one key, 25% over 30 more, the rest over 500K keys), a selection, and
    one key equal to INT64_MIN: with hot keys (QDB_GPU_HOT_MIN_ROWS lowers
    their row threshold) a sample names the frequent keys, the scatter
    aggregates their rows in LDS and one record per hot key joins the
    update; results equal the oracle with and without, in an empty table
    (pending runs) and then in a table that already holds groups (merged)."""
    monkeypatch.setenv("QDB_GPU_BUCKETED_MIN_ROWS", str(1 << 20))
    monkeypatch.setenv("QDB_GPU_HOT_MIN_ROWS", str(1 << 20))
    monkeypatch.setenv("QDB_GPU_HOT_KEYS", hot)
    rng = np.random.default_rng(91)
    n = 1_300_000
    u = rng.random(n)
    key = np.where(u < 0.4, 11, np.where(u < 0.65, rng.integers(100, 131, n), rng.integers(-250_000, 250_000, n))).astype(np.int64)
    key[rng.integers(0, n, 5000)] = np.iinfo(np.int64).min
    vi = rng.integers(-10 ** 12, 10 ** 12, n).astype(np.int64)
    vf = rng.normal(0, 1e3, n)
    sel = (rng.random(n) > 0.1).astype(np.uint8)
    types = [T(SIGNED, 8), T(SIGNED, 8), T(FLOAT, 8)]
    case = AggCase(types, [0], [(AGG_COUNT_STAR, -1), (AGG_SUM, 1), (AGG_MIN, 1), (AGG_MIN, 2), (AGG_MAX, 2), (AGG_SUM, 2)])
    b = Batch(types, [key, vi, vf], None, sel)
    got = sorted(result_rows(gpu_backend.aggregate(case, [b, b], capacity=1 << 21)), key=sort_key)
    want = sorted(result_rows(OracleBackend(fsum_mode=2).aggregate(case, [b, b])), key=sort_key)
    same(got, want)
