"""Served form of the path: a request loop over HBM-resident tables with
continuous batching on several streams.

Tables: `tenants` copies (seeded apart) of the three query shapes' tables
stay in HBM for the whole stream -- the filtered lineitem group-by (6M rows),
the equi join's probe table (60M rows) with its build side filtered and
hashed ONCE, and the high-cardinality group-by (100M rows) -- about 3.6 GB
per tenant. A request is one shape over a row range of one tenant's table.

Continuous batching: requests are admitted as they arrive while the
estimated hash-table and intermediate bytes of everything admitted stay
under the HBM budget. Admitted requests wait for a worker; the workers are
threads that each issue on their own HIP stream (qdb_gpu_set_thread_stream),
so their work overlaps on the device. Whenever a worker is free it takes a
unit of the waiting requests: every waiting group-by of the oldest
request's shape (any tenant) as ONE batched launch (qdb_gpu_agg_update_
batched) with one synchronization for their finishes, or every waiting join
of that tenant as ONE replay of the stream's captured probe pipeline
(qdb_gpu_join_probe_detached) against the resident build table. Requests
that arrive meanwhile are admitted at once and dispatched as soon as a
worker frees up; a unit's memory is released when it completes.

The plain loop runs one request at a time on the library stream, each as
its own query (group-by: set-up, update, finish; join: a probe of the
resident table). Integer and decimal results are exact in both, so a
request's result is the same alone and in any mix.
"""
from __future__ import annotations

import concurrent.futures as cf
import itertools
import math
import threading
import time
from collections import deque
from dataclasses import dataclass
from typing import Any, List

import numpy as np

from . import abi, lib
from . import workloads as W
from .host import (AggregateKernels, DeviceBuffer, FilterKernel, RowJoinKernels, TAggregateProcessor, TRowJoinProcessor,
                   free_pair_buffer, set_thread_stream)

SHAPES = ("lineitem", "join", "groups")
TENANT_SEED_STRIDE = 7919


@dataclass
class Request:
    rid: int
    shape: str
    r0: int
    r1: int
    arrival: float                 # seconds after the stream starts
    tenant: int = 0
    start: float = math.nan
    done: float = math.nan
    result: Any = None

    @property
    def latency(self) -> float:
        return self.done - self.arrival


class Tenant:
    """One copy of the three shapes' tables, and the join's build side
    (filtered, appended and hashed once)."""

    def __init__(self, k: int, res: "Resident", lineitem_rows: int, build_rows: int, probe_rows: int, group_rows: int,
                 distinct: int, cdf_buf):
        off = k * TENANT_SEED_STRIDE
        self.li = W.LineitemDevice(lineitem_rows, seed=W.SEED_LINEITEM + off)
        self.gd = W.GroupsDevice(group_rows, distinct, seed=W.SEED_GROUPS + off, cdf=res.cdf)
        self.jd = W.JoinDevice(build_rows, probe_rows, seed=W.SEED_JOIN + off)
        self._sel = DeviceBuffer(build_rows)
        import ctypes
        view = type(self.jd.build.rs).from_buffer_copy(self.jd.build.rs)
        view.Selection = self._sel.ptr
        res.pred.fn(ctypes.byref(view), None, None, None)
        self.jd.build.set_selection_buffer(self._sel)
        self.join = TRowJoinProcessor(res.join_kern, build_is_left=False)
        self.join.append(self.jd.build)
        self.join.build()


class Resident:
    """Every tenant's tables, generated in HBM once, and one compiled kernel
    set per shape shared by all tenants."""

    def __init__(self, lineitem_rows: int = 6_000_000, build_rows: int = 1_500_000, probe_rows: int = 60_000_000,
                 group_rows: int = 100_000_000, distinct: int = 10_000_000, tenants: int = 1):
        self.li_kern = AggregateKernels(W.lineitem_fused_spec())
        self.g_kern = AggregateKernels(W.groups_spec())
        self.join_kern = RowJoinKernels(W.join_spec(fused=True))
        self.pred = FilterKernel(W.BUILD_TYPES, *W.join_build_predicate())
        self.distinct = distinct
        self.cdf = W.zipf1_cdf(distinct)
        self.sizes = (lineitem_rows, build_rows, probe_rows, group_rows)
        self.tenants = [Tenant(k, self, lineitem_rows, build_rows, probe_rows, group_rows, distinct, None) for k in range(tenants)]
        lib().qdb_gpu_synchronize()
        t0 = self.tenants[0]
        self.li, self.gd, self.jd, self.join = t0.li, t0.gd, t0.jd, t0.join

    def rows(self, shape: str) -> int:
        return {"lineitem": self.li.rows, "join": self.jd.probe_rows, "groups": self.gd.rows}[shape]

    def table_bytes(self) -> int:
        li, b, p, g = self.sizes
        return len(self.tenants) * (li * 38 + b * 12 + p * 16 + g * 24)

    def close(self):
        for t in self.tenants:
            t.join.close()


def make_stream(res: Resident, n: int, qps: float, seed: int = 0xC4, frac_lo: float = 0.01, frac_hi: float = 1.0) -> List[Request]:
    """n requests with Poisson arrivals at `qps`; each one shape (uniform)
    on one tenant (uniform) over a row range whose size is a log-uniform
    fraction of its table."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / qps, n))
    out = []
    for i in range(n):
        shape = SHAPES[int(rng.integers(0, len(SHAPES)))]
        tenant = int(rng.integers(0, len(res.tenants)))
        rows = res.rows(shape)
        size = max(1, int(math.exp(rng.uniform(math.log(frac_lo), math.log(frac_hi))) * rows))
        r0 = int(rng.integers(0, rows - size + 1))
        out.append(Request(i, shape, r0, r0 + size, float(arrivals[i]), tenant))
    return out


def _capacity(rows: int, distinct: int) -> int:
    c = 8
    while c < 2 * min(rows, distinct):
        c *= 2
    return c


class Server:
    """Runs request streams over the resident tables, with continuous
    batching on `workers` streams or one request at a time."""

    def __init__(self, res: Resident, hbm_budget: float = 200e9, full_results: bool = False, workers: int = 3,
                 max_unit: int = 16):
        self.res = res
        self.budget = hbm_budget - res.table_bytes()
        self.full = full_results
        self.workers = workers
        self.max_unit = max_unit
        self._pool = None

    # --- footprints --------------------------------------------------------

    def footprint(self, r: Request) -> int:
        """Estimated device bytes a request adds while it runs: its hash table
        (128-byte records at load 1/2) and intermediates (staged rows, pairs)."""
        n = r.r1 - r.r0
        if r.shape == "join":
            return 16 * n + 4096
        distinct = 6 if r.shape == "lineitem" else self.res.distinct
        return _capacity(n, distinct) * (256 if r.shape == "lineitem" else 128) + 64 * n

    # --- one request, one unit ---------------------------------------------

    def _kern_table(self, r: Request):
        t = self.res.tenants[r.tenant]
        if r.shape == "lineitem":
            return self.res.li_kern, t.li, 6
        return self.res.g_kern, t.gd, self.res.distinct

    def _agg_result(self, shape: str, proc: TAggregateProcessor):
        if shape == "lineitem":
            return W.lineitem_finalize_on_device(proc, proc.kernels)
        if self.full:
            cols = proc.finish()
            return sorted(zip(*[c[0].tolist() for c in cols]))
        return proc.finish_in_hbm()

    def _agg_results_batched(self, shape: str, procs: List[TAggregateProcessor]):
        """The results of every processor, finished with one synchronization."""
        if shape == "lineitem":
            return [W.lineitem_finalize(cols) for cols in TAggregateProcessor.finish_batched(procs)]
        if self.full:
            return [sorted(zip(*[c[0].tolist() for c in cols])) for cols in TAggregateProcessor.finish_batched(procs)]
        return TAggregateProcessor.finish_in_hbm_batched(procs)

    def _join_plain(self, r: Request):
        join = self.res.tenants[r.tenant].join
        jd = self.res.tenants[r.tenant].jd
        before = join.fused_results()
        view = jd.probe.view(r.r0, r.r1)  # alive across the call
        pairs = lib().qdb_gpu_join_processor_probe_only(join.p, view.ptr)
        if pairs < 0:
            raise RuntimeError("join probe failed")
        after = join.fused_results()
        return (int(pairs), after[0]["value"] - before[0]["value"], after[1]["value"] - before[1]["value"])

    def run_plain(self, r: Request, clock) -> None:
        """One request as its own query."""
        r.start = clock()
        if r.shape == "join":
            r.result = self._join_plain(r)
        else:
            kern, table, distinct = self._kern_table(r)
            p = TAggregateProcessor(kern, initial_capacity=_capacity(r.r1 - r.r0, distinct))
            p.add(table.rowset.view(r.r0, r.r1))
            r.result = self._agg_result(r.shape, p)
            p.close()
        r.done = clock()

    def run_unit(self, unit: List[Request], clock) -> None:
        """One unit on the calling worker's stream: group-bys of one shape
        in one batched launch, or joins of one tenant in one replay of the
        stream's probe pipeline."""
        for r in unit:
            r.start = clock()
        shape = unit[0].shape
        if shape == "join":
            t = self.res.tenants[unit[0].tenant]
            views = [t.jd.probe.view(r.r0, r.r1) for r in unit]
            pbs = [abi.QdbPairBuffer() for _ in unit]
            try:
                out = t.join.probe_detached(views, pbs)
            finally:
                for pb in pbs:
                    free_pair_buffer(pb)
            done = clock()
            for r, (n, fused) in zip(unit, out):
                r.result = (n, fused[0]["value"], fused[1]["value"])
                r.done = done
            return
        procs, views = [], []
        for r in unit:
            kern, table, distinct = self._kern_table(r)
            procs.append(TAggregateProcessor(kern, initial_capacity=_capacity(r.r1 - r.r0, distinct)))
            views.append(table.rowset.view(r.r0, r.r1))
        try:
            TAggregateProcessor.add_batched(procs, views)
            results = self._agg_results_batched(shape, procs)
        finally:
            for p in procs:
                p.close()
        done = clock()
        for r, res in zip(unit, results):
            r.result = res
            r.done = done

    # --- the loops ----------------------------------------------------------

    def _worker_pool(self):
        if self._pool is None:
            ids = itertools.count(1)
            lock = threading.Lock()

            def init():
                with lock:
                    k = next(ids)
                set_thread_stream(k)

            self._pool = cf.ThreadPoolExecutor(max_workers=self.workers, initializer=init)
        return self._pool

    def _take_unit(self, queue: List[Request]) -> List[Request]:
        """The oldest waiting request and every waiting request that can
        share its launch (same group-by shape; joins: same tenant)."""
        head = queue[0]
        if head.shape == "join":
            same = [r for r in queue if r.shape == "join" and r.tenant == head.tenant]
        else:
            same = [r for r in queue if r.shape == head.shape]
        unit = same[:self.max_unit]
        ids = {id(r) for r in unit}
        queue[:] = [r for r in queue if id(r) not in ids]
        return unit

    def serve(self, stream: List[Request], batched: bool = True) -> List[Request]:
        """Plays the stream in real time."""
        pending = deque(sorted(stream, key=lambda r: r.arrival))
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0  # noqa: E731
        if not batched:
            while pending:
                now = clock()
                if pending[0].arrival > now:
                    time.sleep(min(pending[0].arrival - now, 0.0005))
                    continue
                self.run_plain(pending.popleft(), clock)
            lib().qdb_gpu_synchronize()
            return stream
        pool = self._worker_pool()
        queue: List[Request] = []
        inflight = {}
        mem = 0
        while pending or queue or inflight:
            now = clock()
            # Admission: every arrived request while the memory budget holds.
            while pending and pending[0].arrival <= now:
                f = self.footprint(pending[0])
                if (queue or inflight) and mem + f > self.budget:
                    break
                r = pending.popleft()
                queue.append(r)
                mem += f
            # Dispatch: a unit to every free worker.
            while queue and len(inflight) < self.workers:
                unit = self._take_unit(queue)
                inflight[pool.submit(self.run_unit, unit, clock)] = unit
            # Wait for a completion or the next arrival. A request that has
            # arrived but does not fit the budget waits for a completion: the
            # loop must not spin, since a spinning thread holds the GIL the
            # workers need after each library call.
            wait_s = 0.0005
            if pending:
                ahead = pending[0].arrival - clock()
                wait_s = min(wait_s, ahead) if ahead > 0 else 0.05
            if inflight:
                done, _ = cf.wait(list(inflight), timeout=wait_s, return_when=cf.FIRST_COMPLETED)
                for fut in done:
                    unit = inflight.pop(fut)
                    fut.result()  # raises a worker's error here
                    mem -= sum(self.footprint(r) for r in unit)
            elif wait_s > 0:
                time.sleep(wait_s)
        return stream

    def close(self):
        if self._pool is not None:
            self._pool.shutdown(wait=True)
            self._pool = None


def summarize(stream: List[Request]) -> dict:
    lat = np.array([r.latency for r in stream]) * 1e3
    span = max(r.done for r in stream) - min(r.arrival for r in stream)
    return {"requests": len(stream), "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "qps": len(stream) / span if span > 0 else float("inf")}


def fresh(stream: List[Request]) -> List[Request]:
    """The same requests with no results (to replay a stream)."""
    return [Request(r.rid, r.shape, r.r0, r.r1, r.arrival, r.tenant) for r in stream]
