"""The batched form against per-instance plain runs: many joins built with one
synchronization and probed in one persistent launch from a device work queue
give, instance by instance, the pairs and fused results of separate
Append/Build/Probe calls (bit-exact), and one instance equals the oracle."""
import ctypes as C

import numpy as np
import pytest

import cases
from cases import AGG_COUNT_STAR, AGG_MIN, AGG_SUM, Batch, GpuBackend, JoinCase
from helpers import SIGNED, T

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu_backend():
    return GpuBackend()


def sorted_pairs(p):
    return p[np.lexsort((p[:, 1], p[:, 0]))]


def join_instances(rng):
    """(build keys, build values, build selection, probe keys, probe values,
    probe selection) per instance: dense unique keys, duplicate keys (hash
    table), a wide unique key range (hash table), an empty probe, a probe
    selection, a one-row build, and a large dense instance."""
    out = []
    b = rng.permutation(5000).astype(np.int64) + 1
    out.append((b, None, rng.integers(-100, 5100, 100_000)))
    b = np.repeat(np.arange(3000, dtype=np.int64), rng.integers(1, 4, 3000))
    out.append((b[rng.permutation(len(b))], None, rng.integers(-50, 3050, 60_000)))
    b = rng.choice(np.iinfo(np.int64).max // 4, 20_000, replace=False).astype(np.int64) - (1 << 60)
    out.append((b, None, np.concatenate([b[rng.integers(0, len(b), 30_000)], rng.integers(-10 ** 18, 10 ** 18, 30_000)])))
    out.append((rng.permutation(700).astype(np.int64), None, np.empty(0, np.int64)))
    b = rng.permutation(40_000).astype(np.int64)
    out.append((b, (rng.random(len(b)) > 0.4).astype(np.uint8), rng.integers(0, 40_000, 250_000)))
    out.append((np.array([42], np.int64), None, rng.integers(40, 45, 5_000)))
    b = rng.permutation(200_000).astype(np.int64) + 7
    out.append((b, None, rng.integers(0, 200_010, 2_000_000)))
    return out


def make_batches(rng, inst, probe_sel_every=4):
    types_b = [T(SIGNED, 8), T(SIGNED, 8)]
    types_p = [T(SIGNED, 8), T(SIGNED, 8)]
    res = []
    for k, (bk, bsel, pk) in enumerate(inst):
        bv = rng.integers(-10 ** 12, 10 ** 12, len(bk)).astype(np.int64)
        pk = np.asarray(pk, np.int64)
        pv = rng.integers(-10 ** 9, 10 ** 9, len(pk)).astype(np.int64)
        psel = (rng.random(len(pk)) > 0.5).astype(np.uint8) if (k % probe_sel_every == 3 and len(pk)) else None
        res.append((Batch(types_b, [bk, bv], None, bsel), Batch(types_p, [pk, pv], None, psel)))
    return types_b, types_p, res


@pytest.mark.parametrize("join_type", [cases.JOIN_INNER, cases.JOIN_LEFT_OUTER])
def test_batched_join_equals_plain(gpu_backend, join_type):
    rng = np.random.default_rng(31 + join_type)
    inst = join_instances(rng)
    types_b, types_p, batches = make_batches(rng, inst)
    case = JoinCase(types_b, types_p, [0], [0], join_type=join_type, probe_is_left=True,
                    fused=[(AGG_COUNT_STAR, -1), (AGG_SUM, 1), (AGG_SUM, -3), (AGG_MIN, -3)])
    devs = [(gpu_backend.dev(b), gpu_backend.dev(p)) for b, p in batches]
    H = gpu_backend.H
    kern, first = gpu_backend.join_processor(case)
    first.close()
    # Plain: one processor per instance, Append / Build / Probe each.
    plain = []
    for db, dp in devs:
        proc = H.TRowJoinProcessor(kern)
        proc.append(db)
        proc.build()
        n = proc.probe_only(dp)
        plain.append((n, sorted_pairs(proc.pairs()), proc.fused_results(), proc.pairs_checksum()))
        proc.close()
    # Batched: every build with one synchronization, every probe in one launch.
    procs = [H.TRowJoinProcessor(kern) for _ in devs]
    for p, (db, _) in zip(procs, devs):
        p.append(db)
    H.TRowJoinProcessor.build_batched(procs)
    for rnd in range(2):
        H.TRowJoinProcessor.probe_batched(procs, [dp for _, dp in devs])
        for k, p in enumerate(procs):
            n, pairs, fused, checksum = plain[k]
            assert p.pair_count() == n, f"instance {k}, round {rnd}: {p.pair_count()} pairs vs {n}"
            assert np.array_equal(sorted_pairs(p.pairs()), pairs), f"instance {k}, round {rnd}: pairs differ"
            assert p.pairs_checksum() == checksum
            # Fused results accumulate over the probes since Init.
            got = p.fused_results()
            for f, (g, w) in enumerate(zip(got, fused)):
                if w["count"] == 0:
                    assert g["count"] == 0 and g["valid"] == w["valid"]
                    continue
                # fused[3] is a MIN: the same after a second probe; the
                # counts and sums double.
                want_val = w["value"] if f == 3 else w["value"] * (rnd + 1)
                assert g["count"] == w["count"] * (rnd + 1), f"instance {k}, round {rnd}: fused count"
                assert g["value"] == want_val, f"instance {k}, round {rnd}: fused {f}: {g} vs {w}"
    # One instance against the oracle.
    b, p = batches[1]
    oj = O.OracleJoin(case.to_c())
    oj.append(O.HostRowSet(types_b, b.columns, selection=b.selection))
    oj.build()
    want = oj.probe(O.HostRowSet(types_p, p.columns, selection=p.selection))
    assert np.array_equal(plain[1][1], sorted_pairs(want))
    for p in procs:
        p.close()


def test_batched_join_rejects_mixed_joins(gpu_backend):
    """Processors of two compiled joins cannot share a batched launch."""
    types = [T(SIGNED, 8), T(SIGNED, 8)]
    c1 = JoinCase(types, types, [0], [0], fused=[(AGG_COUNT_STAR, -1)])
    c2 = JoinCase(types, types, [0], [0], fused=[(AGG_SUM, 1)])
    p1 = gpu_backend.join_processor(c1)[1]
    p2 = gpu_backend.join_processor(c2)[1]
    b = gpu_backend.dev(Batch(types, [np.arange(10, dtype=np.int64), np.arange(10, dtype=np.int64)]))
    p1.append(b)
    p2.append(b)
    with pytest.raises(RuntimeError, match="different joins"):
        gpu_backend.H.TRowJoinProcessor.build_batched([p1, p2])
    p1.close()
    p2.close()


def pairs_of(gpu_backend, pb):
    n = int(pb.Count)
    out = np.empty(2 * n, dtype=np.int64)
    if n:
        gpu_backend.q.check(gpu_backend.q.lib().qdb_gpu_memcpy_d2h(out.ctypes.data, pb.Data, out.nbytes), "d2h")
    return sorted_pairs(out.reshape(n, 2))


@pytest.mark.parametrize("dup", [False, True], ids=["dense", "hash"])
def test_detached_probe_equals_plain(gpu_backend, dup):
    """Many probe batches against ONE built join through the captured probe
    pipeline (qdb_gpu_join_probe_detached): each batch's pairs, count and
    fused results equal a plain probe of that batch alone; the graph is
    replayed several times with plain probes between the replays, with and
    without pair buffers, and from two threads on their own streams at once."""
    import threading
    H = gpu_backend.H
    abi = gpu_backend.q.abi
    rng = np.random.default_rng(5 + dup)
    types_b = [T(SIGNED, 8), T(SIGNED, 8)]
    types_p = [T(SIGNED, 8), T(SIGNED, 8)]
    nb = 60_000
    bk = rng.permutation(nb).astype(np.int64) + 3
    if dup:
        bk = np.concatenate([bk, bk[: nb // 3]])
    bv = rng.integers(-10 ** 12, 10 ** 12, len(bk)).astype(np.int64)
    case = JoinCase(types_b, types_p, [0], [0], fused=[(AGG_COUNT_STAR, -1), (AGG_SUM, 1), (AGG_SUM, -3)])
    kern, proc = gpu_backend.join_processor(case)
    proc.append(gpu_backend.dev(Batch(types_b, [bk, bv])))
    proc.build()
    sizes = [0, 1, 700, 50_000, 333_333, 1_200_000, 64, 90_000]
    probes = []
    for k, n in enumerate(sizes):
        pk = rng.integers(-1000, nb + 1000, n).astype(np.int64)
        pv = rng.integers(-10 ** 9, 10 ** 9, n).astype(np.int64)
        sel = (rng.random(n) > 0.3).astype(np.uint8) if k == 4 else None
        probes.append(gpu_backend.dev(Batch(types_p, [pk, pv], None, sel)))
    # Plain: one probe per batch on a processor of its own.
    want = []
    for d in probes:
        p = H.TRowJoinProcessor(kern)
        p.append(gpu_backend.dev(Batch(types_b, [bk, bv])))
        p.build()
        n = p.probe_only(d)
        f = p.fused_results()
        want.append((n, [x["value"] for x in f], [x["count"] for x in f], [x["valid"] for x in f], sorted_pairs(p.pairs())))
        p.close()

    def check(out, pbs):
        for k, (n, fused) in enumerate(out):
            wn, wv, wc, wvalid, wp = want[k]
            assert n == wn, f"batch {k}: {n} pairs vs {wn}"
            assert [x["value"] for x in fused] == wv and [x["count"] for x in fused] == wc, f"batch {k}: fused"
            assert [x["valid"] for x in fused] == wvalid
            if pbs is not None:
                assert np.array_equal(pairs_of(gpu_backend, pbs[k]), wp), f"batch {k}: pairs"

    for rnd in range(3):
        pbs = [abi.QdbPairBuffer() for _ in probes]
        check(proc.probe_detached(probes, pbs), pbs)
        for pb in pbs:
            H.free_pair_buffer(pb)
        check(proc.probe_detached(probes), None)
        # Eager work between replays.
        assert proc.probe_only(probes[3]) == want[3][0]
    # Two threads, each on its own stream, probing the same table at once.
    errors = []

    def worker(idx):
        try:
            H.set_thread_stream(idx)
            for _ in range(3):
                pbs = [abi.QdbPairBuffer() for _ in probes]
                check(proc.probe_detached(probes, pbs), pbs)
                for pb in pbs:
                    H.free_pair_buffer(pb)
        except Exception as e:  # reported below
            errors.append(repr(e))

    ts = [threading.Thread(target=worker, args=(i,)) for i in (1, 2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors
    proc.close()


def test_batched_finish_small_tables_equal_plain(gpu_backend):
    """A batched finish of many tables with one-word keys: the tables of at
    most 65,536 slots and 16,384 groups are finished by one launch (one
    workgroup per table lists its occupied slots in LDS), the others as by
    their own finish; every output equals that table's plain update and
    finish (as sorted rows: concurrent inserts place groups in slots in any
    order), and one equals the oracle. Some instances
    hold the key equal to the slot table's empty marker (INT64_MIN, kept in
    the extra record), some grow from an 8-slot table, one is empty."""
    import qumirdb_amd as q
    from cases import AGG_MAX, OracleBackend, AggCase
    from helpers import FLOAT, result_rows, sort_key
    rng = np.random.default_rng(77)
    types = [T(SIGNED, 8), T(SIGNED, 8), T(FLOAT, 8)]
    case = AggCase(types, [0], [(AGG_COUNT_STAR, -1), (AGG_SUM, 1), (AGG_MIN, 2), (AGG_MAX, 2), (AGG_SUM, 2)])
    specs = [(3000, 40), (1, 1), (0, 1), (50_000, 9000), (200_000, 15_000), (120_000, 30_000), (80_000, 300),
             (10_000, 10_000), (64_000, 5), (2_000, 2_000), (300_000, 60_000), (40_000, 4_000)]
    batches, caps = [], []
    for k, (n, card) in enumerate(specs):
        keys = rng.integers(-card, card, n).astype(np.int64) * 7919
        if k % 3 == 1 and n:
            keys[rng.integers(0, n, max(1, n // 10))] = np.iinfo(np.int64).min
        vi = rng.integers(-10 ** 12, 10 ** 12, n).astype(np.int64)
        vf = rng.normal(0, 1e3, n)
        batches.append(Batch(types, [keys, vi, vf]))
        caps.append(8 if k % 4 == 0 else 1 << max(3, int(2 * min(n, 2 * card)).bit_length()))
    kern = q.AggregateKernels.__new__(q.AggregateKernels)
    kern.spec = None
    kern.c_spec = gpu_backend.cast(case.to_c())
    kern.k = q.abi.QdbAggregateKernels()
    q.check(q.lib().qdb_gpu_compile_aggregate(C.byref(kern.c_spec), C.byref(kern.k)), "compile_aggregate")
    kern.out_types = [gpu_backend.H.ValueType(t.Kind, t.Width, t.Scale, bool(t.Nullable))
                      for t in kern.k.OutputTypes[: kern.k.NumOutputColumns]]
    devs = [gpu_backend.dev(b) for b in batches]
    plain = []
    for d, cap in zip(devs, caps):
        p = q.TAggregateProcessor(kern, initial_capacity=cap)
        p.add(d)
        plain.append(sorted(result_rows(p.finish()), key=sort_key))
        p.close()
    procs = [q.TAggregateProcessor(kern, initial_capacity=cap) for cap in caps]
    for p, d in zip(procs, devs):
        p.add(d)
    finished = q.TAggregateProcessor.finish_batched(procs)
    for i, p in enumerate(procs):
        p.close()
        assert sorted(result_rows(finished[i]), key=sort_key) == plain[i], f"instance {i} ({specs[i]})"
    want = sorted(result_rows(OracleBackend(fsum_mode=2).aggregate(case, [batches[1 + 3]])), key=sort_key)
    assert plain[4] == want
    q.lib().qdb_gpu_release_slot(q.abi.SLOT_AGGREGATE, kern.k.Slot)
