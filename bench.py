"""Throughput of the MI355X hash-join / hash-aggregate path.

The headline workload is the equi hash join with a filtered build: build
1.5M rows (int64 key = a permutation of 1..1.5M, segment uniform 0..4,
filter segment = 1), probe 60M rows (int64 fk uniform over 1..1.5M,
DECIMAL(15,2) value); outputs COUNT(*), exact SUM(value) and the matched
(probe_row_id, build_row_id) pairs. One step is one whole query on inputs
already resident in HBM: build filter -> append -> table build -> probe
(pairs + fused COUNT/SUM). `value` is (build rows + probe rows) / second.

The other shapes are measured in the same run (unless --workload join):
the filtered low-cardinality group-by over 6M lineitem rows (fused filter +
project + aggregate), the lineitem projection alone over 60M rows, the
high-cardinality Zipf group-by over 100M rows, and the batched and served
forms.

roofline: algorithmic bytes of the dominant kernel (the probe: 16 B per
probe row read + 16 B per pair written) over its average duration, timed
with HIP events on the library's stream. cpu_baseline: the CPU oracle (a
restatement of the reference algorithm) over the whole join, its probe in
262,144-row morsels on the host's threads, and on one thread beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["join", "lineitem", "project", "groups", "batched", "served", "all"],
                    default="all")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run a plain device copy of MIB MiB (3 launches), the known-size kernel that "
                         "scales rocprof FETCH_SIZE / WRITE_SIZE readings")
    return ap.parse_args()


def timing(family):
    import ctypes
    from qumirdb_amd import lib
    n = ctypes.c_int64()
    ms = lib().qdb_gpu_timing_avg_ms(family.encode(), ctypes.byref(n))
    return ms, n.value


def timing_per_step(family, steps):
    """Milliseconds per step of every region of `family` (a step may time
    several regions of one family)."""
    ms, n = timing(family)
    return ms * n / steps if steps else 0.0


def run_steps(step, steps):
    """Times `steps` steps on the wall clock with the library's HIP-event
    regions off (the throughput figure), then `steps` more with them on (the
    per-kernel durations, read with timing() afterwards; recording the events
    costs time of its own). Returns (ms per step, ms per step with events)."""
    from qumirdb_amd import lib
    lib().qdb_gpu_synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    lib().qdb_gpu_synchronize()
    t1 = time.perf_counter()
    lib().qdb_gpu_timing_reset()
    lib().qdb_gpu_timing_enable(1)
    t2 = time.perf_counter()
    for _ in range(steps):
        step()
    lib().qdb_gpu_synchronize()
    t3 = time.perf_counter()
    lib().qdb_gpu_timing_enable(0)
    return (t1 - t0) * 1e3 / steps, (t3 - t2) * 1e3 / steps


def bench_join(steps, warmup, build_rows=1_500_000, probe_rows=60_000_000, table_mode=0, jd=None):
    """table_mode 0: the build picks its table (dense direct table for these
    keys); 1: the hash table."""
    import ctypes
    from qumirdb_amd import lib, workloads as W
    from qumirdb_amd.host import DeviceBuffer, FilterKernel, RowJoinKernels, TRowJoinProcessor

    jd = jd or W.JoinDevice(build_rows, probe_rows)
    pred = FilterKernel(W.BUILD_TYPES, *W.join_build_predicate())
    kern = RowJoinKernels(W.join_spec(fused=True, table_mode=table_mode))
    sel = DeviceBuffer(build_rows)
    results = {}

    # the build-side filter writes the selection into the build batch
    view = type(jd.build.rs).from_buffer_copy(jd.build.rs)
    view.Selection = sel.ptr
    view_ref = ctypes.byref(view)

    def step():
        pred.fn(view_ref, None, None, None)
        jd.build.set_selection_buffer(sel)
        proc = TRowJoinProcessor(kern, build_is_left=False)
        proc.append(jd.build)
        proc.build()
        n = lib().qdb_gpu_join_processor_probe_only(proc.p, jd.probe.ptr)
        f = proc.fused_results()
        results["pairs"], results["count"], results["sum"] = n, f[0]["value"], f[1]["value"]
        proc.close()

    for _ in range(warmup):
        step()
    ms_per_step, ms_events = run_steps(step, steps)
    probe_ms, launches = timing("join_probe")
    build_ms, nb = timing("join_build")
    place_ms, npl = timing("join_build_place")
    build_ms = (build_ms * nb + place_ms * npl) / steps
    pairs = results["pairs"]
    probe_bytes = probe_rows * 16 + pairs * 16
    achieved = probe_bytes / (probe_ms * 1e-3) / 1e9 if probe_ms else 0.0
    lib().qdb_gpu_timing_reset()
    return {
        "rows": build_rows + probe_rows, "ms_per_step": ms_per_step, "ms_per_step_with_events": ms_events,
        "rows_per_s": (build_rows + probe_rows) / (ms_per_step * 1e-3),
        "probe_kernel_ms": probe_ms, "probe_launches": launches, "build_kernel_ms": build_ms,
        "pairs": pairs, "count": results["count"], "sum_value_unscaled": str(results["sum"]),
        "algorithmic_bytes_per_probe": probe_bytes, "achieved_gbs": achieved,
        "table": "hash" if table_mode == 1 else "auto",
    }, jd


def bench_lineitem(steps, warmup, rows=6_000_000):
    from qumirdb_amd import lib, workloads as W
    from qumirdb_amd.host import AggregateKernels, TAggregateProcessor

    dev = W.LineitemDevice(rows)
    kern = AggregateKernels(W.lineitem_fused_spec())
    out = {}

    def step():
        p = TAggregateProcessor(kern, initial_capacity=8)
        p.add(dev.rowset)
        out["rows"] = W.lineitem_finalize_on_device(p, kern)
        p.close()

    for _ in range(warmup):
        step()
    ms, ms_events = run_steps(step, steps)
    k_ms = timing_per_step("agg_update", steps)
    lib().qdb_gpu_timing_reset()
    nbytes = rows * (1 + 1 + 4 + 8 * 4)
    return {"rows": rows, "ms_per_step": ms, "ms_per_step_with_events": ms_events, "rows_per_s": rows / (ms * 1e-3), "kernel_ms": k_ms,
            "algorithmic_bytes": nbytes, "achieved_gbs": nbytes / (k_ms * 1e-3) / 1e9 if k_ms else 0.0,
            "groups": len(out["rows"])}


def bench_project(steps, warmup, rows=60_000_000):
    """The lineitem projection price * (1 - discount) and price * (1 -
    discount) * (1 + tax) (two i128 DECIMAL outputs) over rows in HBM, into
    output columns allocated once; 56 algorithmic bytes per row (three
    8-byte inputs, two 16-byte outputs)."""
    import ctypes
    from qumirdb_amd import lib, workloads as W
    from qumirdb_amd.host import ProjectKernel

    dev = W.LineitemDevice(rows)
    kern = ProjectKernel(W.LINEITEM_TYPES, [W.lineitem_disc_price(), W.lineitem_charge()])
    bufs, ptrs = kern(dev.rowset)

    def step():
        kern.fn(ctypes.byref(dev.rowset.rs), ptrs, None, None, None)

    for _ in range(warmup):
        step()
    ms, ms_events = run_steps(step, steps)
    err = lib().qdb_gpu_last_error()
    if err:
        raise RuntimeError(f"project: {err.decode()}")
    k_ms = timing_per_step("project", steps)
    lib().qdb_gpu_timing_reset()
    del bufs
    nbytes = rows * (8 * 3 + 16 * 2)
    return {"rows": rows, "ms_per_step": ms, "ms_per_step_with_events": ms_events, "rows_per_s": rows / (ms * 1e-3),
            "kernel_ms": k_ms, "algorithmic_bytes": nbytes, "achieved_gbs": nbytes / (k_ms * 1e-3) / 1e9 if k_ms else 0.0}


def bench_groups(steps, warmup, rows=100_000_000, distinct=10_000_000):
    from qumirdb_amd import lib, workloads as W
    from qumirdb_amd.host import AggregateKernels, TAggregateProcessor

    dev = W.GroupsDevice(rows, distinct)
    kern = AggregateKernels(W.groups_spec())
    out = {}

    def step():
        p = TAggregateProcessor(kern, initial_capacity=1 << 24)
        p.add(dev.rowset)
        out["groups"] = p.finish_in_hbm()
        p.close()

    # The first warm-up query tunes the shape (the autotuner times its
    # candidate configurations on scratch tables and caches the fastest).
    was_tuning = lib().qdb_gpu_autotune_enable(1)
    for _ in range(warmup):
        step()
    lib().qdb_gpu_autotune_enable(was_tuning)
    tuned = tuned_configs()
    ms, ms_events = run_steps(step, steps)
    k_ms = timing_per_step("agg_update", steps)
    f_ms = timing_per_step("agg_finalize", steps)  # occupancy scan, then the output rows
    lib().qdb_gpu_timing_reset()
    groups = out["groups"]
    nbytes = rows * 24 + groups * 48
    return {"rows": rows, "ms_per_step": ms, "ms_per_step_with_events": ms_events, "rows_per_s": rows / (ms * 1e-3),
            "update_kernel_ms": k_ms,
            "finalize_kernel_ms": f_ms, "groups": groups, "algorithmic_bytes": nbytes,
            "achieved_gbs": nbytes / ((k_ms + f_ms) * 1e-3) / 1e9 if k_ms else 0.0, "autotuned": tuned}


def tuned_configs():
    """The autotuner's cache: the configuration chosen per shape bucket."""
    from qumirdb_amd import abi, lib
    n = lib().qdb_gpu_autotune_entries(None, 0)
    ents = (abi.QdbTuneEntry * max(1, n))()
    lib().qdb_gpu_autotune_entries(ents, n)
    return [{"key_words": e.KeyWords, "card_bucket": e.CardBucket, "rows_bucket": e.RowsBucket,
             "config": ("bucketed, fan-out x%g, partition over <= %d workgroups" % (2.0 ** e.FanoutShift, e.PartBlocks or 2048))
             if e.Path == 1 else f"LDS {e.LdsKiB} KiB x {e.GridPerCu}/CU" + (f", {e.LdsCopies} copies" if e.LdsCopies > 1 else ""),
             "ms": round(e.Ms, 3)} for e in ents[:n]]


def bench_batched(n_groupby=256, n_join=256):
    """The batched form: 512 instances, 256 group-bys (COUNT(*), SUM(i64),
    MIN/MAX/SUM(f64) by an int64 key; rows log-uniform 10K..20M, cardinality
    log-uniform 4..1M, uniform or Zipf(1.1) keys) and 256 equi joins (build
    log-uniform 1K..2M unique int64 keys, probe log-uniform 10K..20M rows;
    COUNT(*), exact SUM(value) and the matched pairs). Plain: whole queries
    one after another (group-by: table set-up, update, finish into HBM; join:
    append, build, probe). Batched: every group-by update in one persistent
    launch and every finish behind one synchronization; every join build
    behind one synchronization and every probe in one persistent launch;
    the joins run on a stream of their own beside the group-bys, and the
    group-by instances on the bucketed path on the library stream and two
    helper streams. Results are compared instance by instance afterwards: group-by outputs
    in full (downloaded and sorted), joins by pair count, fused COUNT/SUM and
    an order-independent checksum of the pairs."""
    import ctypes
    import threading

    import numpy as np

    from qumirdb_amd import lib, workloads as W
    from qumirdb_amd.host import (AggregateKernels, RowJoinKernels, TAggregateProcessor, TRowJoinProcessor,
                                  download_rowset, set_thread_stream)

    inst = W.batched_instances(n_groupby)
    jinst = W.batched_join_instances(n_join)
    dev = W.BatchedDevice(inst)
    jdev = W.BatchedJoinDevice(jinst)
    kern = AggregateKernels(W.groups_spec())
    jk = RowJoinKernels(W.batched_join_spec())

    def capacity(rows, card):
        c = 8
        while c < 2 * min(rows, card):
            c *= 2
        return c

    caps = [capacity(r, c) for r, c, _ in inst]

    def plain():
        outs = []
        for d, cap in zip(dev.devs, caps):
            p = TAggregateProcessor(kern, initial_capacity=cap)
            p.add(d.rowset)
            outs.append(p.finish_raw())
            p.close()
        jps = []
        for jd in jdev.devs:
            jp = TRowJoinProcessor(jk)
            jp.append(jd.build)
            jp.build()
            jp.probe_only(jd.probe)
            jp.fused_results()
            jps.append(jp)
        return outs, jps

    def joins_batched(out):
        # On a stream of its own, beside the group-bys.
        try:
            set_thread_stream(3)
            jps = [TRowJoinProcessor(jk) for _ in jdev.devs]
            for jp, jd in zip(jps, jdev.devs):
                jp.append(jd.build)
            TRowJoinProcessor.build_batched(jps)
            TRowJoinProcessor.probe_batched(jps, [jd.probe for jd in jdev.devs])
            for jp in jps:
                jp.fused_results()
            lib().qdb_gpu_synchronize()
            out.append(jps)
        except Exception as e:  # re-raised by the caller
            out.append(e)

    def batched():
        jout = []
        jt = threading.Thread(target=joins_batched, args=(jout,))
        jt.start()
        procs = [TAggregateProcessor(kern, initial_capacity=cap) for cap in caps]
        TAggregateProcessor.add_batched(procs, [d.rowset for d in dev.devs])
        raw = TAggregateProcessor._finish_batched_raw(procs)
        outs = [raw[i] for i in range(len(procs))]
        for p in procs:
            p.close()
        jt.join()
        if isinstance(jout[0], Exception):
            raise jout[0]
        return outs, jout[0]

    def release(res):
        outs, jps = res
        for o in outs:
            lib().qdb_gpu_rowset_release(ctypes.byref(o))
        for jp in jps:
            jp.close()

    types = kern.out_types

    def digest(res):
        """Every group-by output downloaded and sorted by key (f64 as bits),
        and every join's pair count, fused results and pair checksum; then
        the results are released."""
        outs, jps = res
        gd = []
        for o in outs:
            r = download_rowset(o, types)
            order = np.argsort(np.asarray(r[0][0]), kind="stable")
            cols = []
            for c in range(len(types)):
                v = np.asarray(r[c][0])[order]
                cols.append(v.view(np.int64) if v.dtype.kind == "f" else v)
            gd.append(cols)
        jd = [(jp.pair_count(), jp.fused_results(), jp.pairs_checksum()) for jp in jps]
        release(res)
        return gd, jd

    # Autotuning on: the warm-up runs tune every shape bucket the instances
    # fall in (plain and batched updates then run in the tuned configurations).
    was_tuning = lib().qdb_gpu_autotune_enable(1)
    release(plain())
    release(batched())  # warm-up: kernels compiled, shapes tuned, allocator primed
    lib().qdb_gpu_synchronize()
    # The timed runs: HIP-event regions off (each region's events cost tens
    # of microseconds of host time, and the two forms open different
    # numbers of them).
    lib().qdb_gpu_timing_enable(0)
    t0 = time.perf_counter()
    r_plain = plain()
    lib().qdb_gpu_synchronize()
    t1 = time.perf_counter()
    d_plain = digest(r_plain)
    t2 = time.perf_counter()
    r_batch = batched()
    lib().qdb_gpu_synchronize()
    t3 = time.perf_counter()
    d_batch = digest(r_batch)
    # Kernel durations from one more run of each form with the regions on.
    lib().qdb_gpu_timing_reset()
    lib().qdb_gpu_timing_enable(1)
    release(plain())
    upd_ms, upd_n = timing("agg_update")
    probe_ms, probe_n = timing("join_probe")
    lib().qdb_gpu_timing_reset()
    release(batched())
    bat_ms, _ = timing("agg_batch")
    bprobe_ms, _ = timing("join_probe_batched")
    lib().qdb_gpu_timing_enable(0)
    lib().qdb_gpu_timing_reset()
    lib().qdb_gpu_autotune_enable(was_tuning)
    # Results, instance by instance.
    groups = 0
    for k, (a, b) in enumerate(zip(d_plain[0], d_batch[0])):
        if len(a[0]) != len(b[0]) or not all(np.array_equal(x, y) for x, y in zip(a, b)):
            raise RuntimeError(f"group-by instance {k} differs between plain and batched runs")
        groups += len(a[0])
    pairs = 0
    for k, (a, b) in enumerate(zip(d_plain[1], d_batch[1])):
        if a != b:
            raise RuntimeError(f"join instance {k}: pairs or fused results differ between plain and batched runs")
        pairs += a[0]
    tuned = tuned_configs()
    n = n_groupby + n_join
    rows = sum(r for r, _, _ in inst) + sum(b + p for b, p in jinst)
    return {"instances": n, "groupby_instances": n_groupby, "join_instances": n_join, "rows": rows, "groups": groups,
            "pairs": pairs, "plain_s": t1 - t0, "batched_s": t3 - t2,
            "plain_queries_per_s": n / (t1 - t0), "batched_queries_per_s": n / (t3 - t2),
            "speedup": (t1 - t0) / (t3 - t2), "rows_per_s": rows / (t3 - t2), "ms_per_step": (t3 - t2) * 1e3,
            "plain_update_kernels_ms": upd_ms * upd_n, "batched_update_kernel_ms": bat_ms,
            "plain_probe_kernels_ms": probe_ms * probe_n, "batched_probe_kernel_ms": bprobe_ms,
            "results": "equal (group-by outputs in full; joins: pair counts, fused COUNT/SUM, pair checksums)",
            "autotuned_shapes": tuned}


def bench_served(qps_levels=(50, 200, 1000, 2000, 4000, 8000), n_requests=300):
    """The served form: Poisson request streams at each offered QPS over the
    resident tables (lineitem 6M rows, join build 1.5M / probe 60M rows,
    groups 100M rows; each request one shape over a log-uniform 1%..100%
    row range), continuous batching versus one request at a time, with the
    same results checked request by request."""
    from qumirdb_amd import serve as SV
    res = SV.Resident(tenants=8)
    srv = SV.Server(res)
    warm = SV.make_stream(res, 12, 1e6, seed=1)
    srv.serve(warm, batched=True)
    srv.serve(SV.fresh(warm), batched=False)
    levels = []
    for k, qps in enumerate(qps_levels):
        stream = SV.make_stream(res, n_requests, qps, seed=0xC4 + k)
        plain = srv.serve(SV.fresh(stream), batched=False)
        batched = srv.serve(SV.fresh(stream), batched=True)
        for a, b in zip(plain, batched):
            if a.result != b.result:
                raise RuntimeError(f"request {a.rid} ({a.shape}) differs between plain and batched serving")
        levels.append({"offered_qps": qps, "plain": SV.summarize(plain), "batched": SV.summarize(batched)})
    srv.close()
    res.close()
    return {"levels": levels, "requests_per_level": n_requests,
            "resident_bytes": res.table_bytes()}


def cpu_threads():
    """Host threads for the CPU baseline: the process's CPU share, at most
    OMP_NUM_THREADS (16 per GPU here)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    cap = int(os.environ.get("OMP_NUM_THREADS", "16") or 16)
    return max(1, min(n, cap))


def cpu_join_baseline(build_rows=1_500_000, sample_probe=60_000_000):
    """The oracle's row-table join over the filtered build and `sample_probe`
    probe rows (by default the whole workload): build on one thread, then the
    probe over 262,144-row morsels on the host's threads (the value), and the
    same join on one thread."""
    import numpy as np
    from oracle import oracle as O
    from qumirdb_amd import workloads as W
    key, seg = O.gen_join_build(W.SEED_JOIN, 0, build_rows, build_rows)
    fk, val = O.gen_join_probe(W.SEED_JOIN, 0, sample_probe, build_rows)
    threads = cpu_threads()
    rows = build_rows + sample_probe
    probe = O.HostRowSet(W.PROBE_TYPES, [fk, val])  # kept alive across the calls
    build = O.HostRowSet(W.BUILD_TYPES, [key, seg], selection=(seg == 1).astype(np.uint8))
    times = {}
    for t in (1, threads):
        t0 = time.perf_counter()
        oj = O.OracleJoin(W.join_spec().to_c())
        oj.append(build)
        oj.build()
        n = O.LIB.oracle_join_probe_parallel(oj.h, probe.ptr, t) if t > 1 else O.LIB.oracle_join_probe(oj.h, probe.ptr)
        times[t] = time.perf_counter() - t0
        del oj
    return {"value": rows / times[threads], "unit": "rows/s", "cores": threads, "kind": "port",
            "one_thread_value": rows / times[1],
            "sample": f"build {build_rows} rows (filter segment = 1) + probe {sample_probe} of 60M rows, {n} pairs: "
                      f"{times[threads]:.2f} s on {threads} threads (probe in 262,144-row morsels), "
                      f"{times[1]:.2f} s on one thread"}


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_probe_dense_lean__join_build1.5M_probe60M.md")


def profiled_traffic():
    """HBM bytes per probe launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (counter readings scaled by the copy calibration), or
    None when the summary is absent."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":")[1])
    except OSError:
        pass
    return None


def copy_calibration(mib):
    """A plain copy of `mib` MiB through an elementwise kernel (bitwise NOT,
    3 launches; a same-type copy_ would go to the DMA engine, which no kernel
    counter sees): FETCH_SIZE / WRITE_SIZE of this kernel against its known
    bytes gives the factor that turns the counters into HBM bytes."""
    import torch
    n = mib << 20
    src = torch.ones(n // 8, dtype=torch.int64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        torch.bitwise_not(src, out=dst)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / 3
    del src, dst
    return {"bytes_read": n, "bytes_written": n, "ms": ms, "gbs": 2 * n / (ms * 1e-3) / 1e9}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported in this build "
              "(the path runs on a single MI355X).", file=sys.stderr)
        sys.exit(2)
    import qumirdb_amd
    qumirdb_amd.init_device(0)
    out = {}
    import gc
    from qumirdb_amd import lib as _lib

    def release():
        """Between legs: the previous leg's buffers back to the HIP runtime."""
        gc.collect()
        _lib().qdb_gpu_trim_cache()

    if args.workload in ("join", "all"):
        out["join"], jd = bench_join(args.steps, args.warmup)
        out["join_hash_table"], jd = bench_join(args.steps, args.warmup, table_mode=1, jd=jd)
        del jd
        release()
    if args.workload in ("lineitem", "all"):
        out["lineitem"] = bench_lineitem(args.steps, args.warmup)
        release()
    if args.workload in ("project", "all"):
        out["project"] = bench_project(args.steps, args.warmup)
        release()
    if args.workload in ("groups", "all"):
        # Five untimed queries at least: the first compiles the generated
        # kernels, and the next few still run up to ~1.6x the steady-state
        # time (9.1 ms per query once warm).
        out["groups"] = bench_groups(max(1, args.steps // 2), max(5, args.warmup // 2))
        release()
    if args.workload in ("batched", "all"):
        out["batched"] = bench_batched()
        release()
    if args.workload in ("served", "all"):
        out["served"] = bench_served()
    if args.copy_calibration:
        out["copy_calibration"] = copy_calibration(args.copy_calibration)
    head = out.get("join") or out.get("lineitem") or out.get("project") or out.get("groups") or out.get("batched")
    if head is None:  # served only: its best sustained throughput
        best = max(out["served"]["levels"], key=lambda l: l["batched"]["qps"])["batched"]
        head = {"rows_per_s": best["qps"], "ms_per_step": best["p50_ms"]}
    name = "join" if "join" in out else list(out)[0]
    cpu = None
    if not args.no_cpu and name == "join":
        cpu = cpu_join_baseline()
    roof = None
    if name == "join":
        j = out["join"]
        roof = {"bound": "hbm", "achieved": round(j["achieved_gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(j["achieved_gbs"] / HBM_PEAK_GBS, 4), "traffic": profiled_traffic(),
                "kernel": "k_probe_dense_lean (one pass, fused totals by atomics)", "kernel_ms": round(j["probe_kernel_ms"], 4)}
    import ctypes
    from qumirdb_amd import lib
    limits = (ctypes.c_int64 * 2)()
    lib().qdb_gpu_device_limits(limits, 2)
    line = {
        "metric": "rows/sec + achieved HBM GB/s (% of peak) for agg/join; queries/s and p99 latency",
        "value": head["rows_per_s"], "unit": "rows/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int64", "data": "synthetic (seeded, generated in HBM)",
        "config": {"workload": "equi hash join, filtered build 1.5M rows (segment = 1), probe 60M rows; "
                               "COUNT(*), exact SUM(value), matched pairs" if name == "join" else name,
                   "parallelism": "single GPU"},
        "roofline": roof, "cpu_baseline": cpu,
        "device": {"compute_units": limits[0], "max_lds_per_workgroup": limits[1]},
        "workloads": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()} for k, v in out.items()},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
