"""Served form of the path: a request loop over HBM-resident tables with
continuous batching.

The three query shapes stay resident in HBM for the whole stream: the
filtered lineitem group-by (6M rows), the equi join's probe table (60M rows)
with its build side filtered and hashed ONCE, and the high-cardinality
group-by (100M rows). A request is one shape over a row range of its table.
Requests arrive on their own clock; the server admits every request that
has arrived while the estimated hash-table and intermediate bytes of the
admitted work stay under the HBM budget, and runs them together: group-bys
of one shape share ONE batched launch (qdb_gpu_agg_update_batched) and one
synchronization for their finishes (qdb_gpu_agg_finish_batched), joins
probe the resident build table in turn. Requests that arrive while a round
runs are admitted in the next round, as slots and memory free up. The plain
loop runs one request per round, in arrival order. Integer and decimal
results are exact in both, so a request's result is the same alone and in
any mix.
"""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Any, List, Optional

import numpy as np

from . import lib
from . import workloads as W
from .host import AggregateKernels, DeviceBuffer, FilterKernel, RowJoinKernels, TAggregateProcessor, TRowJoinProcessor

SHAPES = ("lineitem", "join", "groups")


@dataclass
class Request:
    rid: int
    shape: str
    r0: int
    r1: int
    arrival: float                 # seconds after the stream starts
    start: float = math.nan
    done: float = math.nan
    result: Any = None

    @property
    def latency(self) -> float:
        return self.done - self.arrival


class Resident:
    """The tables every request reads, generated in HBM once, and the join's
    build side (filtered, appended and hashed once)."""

    def __init__(self, lineitem_rows: int = 6_000_000, build_rows: int = 1_500_000, probe_rows: int = 60_000_000,
                 group_rows: int = 100_000_000, distinct: int = 10_000_000):
        self.li = W.LineitemDevice(lineitem_rows)
        self.li_kern = AggregateKernels(W.lineitem_fused_spec())
        self.gd = W.GroupsDevice(group_rows, distinct)
        self.g_kern = AggregateKernels(W.groups_spec())
        self.distinct = distinct
        self.jd = W.JoinDevice(build_rows, probe_rows)
        self.join_kern = RowJoinKernels(W.join_spec(fused=True))
        pred = FilterKernel(W.BUILD_TYPES, *W.join_build_predicate())
        self._sel = DeviceBuffer(build_rows)
        import ctypes
        view = type(self.jd.build.rs).from_buffer_copy(self.jd.build.rs)
        view.Selection = self._sel.ptr
        pred.fn(ctypes.byref(view), None, None, None)
        self.jd.build.set_selection_buffer(self._sel)
        self.join = TRowJoinProcessor(self.join_kern, build_is_left=False)
        self.join.append(self.jd.build)
        self.join.build()
        lib().qdb_gpu_synchronize()

    def rows(self, shape: str) -> int:
        return {"lineitem": self.li.rows, "join": self.jd.probe_rows, "groups": self.gd.rows}[shape]

    def table_bytes(self) -> int:
        return self.li.rows * 38 + self.jd.build_rows * 12 + self.jd.probe_rows * 16 + self.gd.rows * 24

    def close(self):
        self.join.close()


def make_stream(res: Resident, n: int, qps: float, seed: int = 0xC4, frac_lo: float = 0.01, frac_hi: float = 1.0) -> List[Request]:
    """n requests with Poisson arrivals at `qps`; each one shape (uniform)
    over a row range whose size is a log-uniform fraction of its table."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / qps, n))
    out = []
    for i in range(n):
        shape = SHAPES[int(rng.integers(0, len(SHAPES)))]
        rows = res.rows(shape)
        size = max(1, int(math.exp(rng.uniform(math.log(frac_lo), math.log(frac_hi))) * rows))
        r0 = int(rng.integers(0, rows - size + 1))
        out.append(Request(i, shape, r0, r0 + size, float(arrivals[i])))
    return out


def _capacity(rows: int, distinct: int) -> int:
    c = 8
    while c < 2 * min(rows, distinct):
        c *= 2
    return c


class Server:
    """Runs request streams over the resident tables, batched or plain."""

    def __init__(self, res: Resident, hbm_budget: float = 200e9, full_results: bool = False):
        self.res = res
        self.budget = hbm_budget - res.table_bytes()
        self.full = full_results

    # --- footprints --------------------------------------------------------

    def footprint(self, r: Request) -> int:
        """Estimated device bytes a request adds while it runs: its hash table
        (128-byte records at load 1/2) and intermediates (spill row lists,
        pairs)."""
        n = r.r1 - r.r0
        if r.shape == "join":
            return 16 * n + 4096
        distinct = 6 if r.shape == "lineitem" else self.res.distinct
        return _capacity(n, distinct) * (256 if r.shape == "lineitem" else 128) + 8 * n

    # --- execution ----------------------------------------------------------

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

    def _agg_procs(self, reqs: List[Request]):
        shape = reqs[0].shape
        kern = self.res.li_kern if shape == "lineitem" else self.res.g_kern
        table = self.res.li if shape == "lineitem" else self.res.gd
        distinct = 6 if shape == "lineitem" else self.res.distinct
        procs = [TAggregateProcessor(kern, initial_capacity=_capacity(r.r1 - r.r0, distinct)) for r in reqs]
        views = [table.rowset.view(r.r0, r.r1) for r in reqs]
        return procs, views

    def _run_join(self, r: Request):
        before = self.res.join.fused_results()
        view = self.res.jd.probe.view(r.r0, r.r1)  # alive across the call
        pairs = lib().qdb_gpu_join_processor_probe_only(self.res.join.p, view.ptr)
        if pairs < 0:
            raise RuntimeError("join probe failed")
        after = self.res.join.fused_results()
        return (int(pairs), after[0]["value"] - before[0]["value"], after[1]["value"] - before[1]["value"])

    def run_round(self, reqs: List[Request], batched: bool, clock) -> None:
        """Runs one admitted round; group-bys of a shape share one batched
        launch when `batched`."""
        for r in reqs:
            r.start = clock()
        for shape in ("lineitem", "groups"):
            group = [r for r in reqs if r.shape == shape]
            if not group:
                continue
            procs, views = self._agg_procs(group)
            if batched:
                TAggregateProcessor.add_batched(procs, views)
                results = self._agg_results_batched(shape, procs)
                done = clock()
                for r, p, res in zip(group, procs, results):
                    r.result = res
                    p.close()
                    r.done = done
                continue
            for p, v in zip(procs, views):
                p.add(v)
            for r, p in zip(group, procs):
                r.result = self._agg_result(shape, p)
                p.close()
                r.done = clock()
        for r in reqs:
            if r.shape == "join":
                r.result = self._run_join(r)
                r.done = clock()

    def serve(self, stream: List[Request], batched: bool = True) -> List[Request]:
        """Plays the stream in real time: a round admits every request that
        has arrived and fits the budget (plain: the oldest one only)."""
        pending = deque(sorted(stream, key=lambda r: r.arrival))
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0  # noqa: E731
        while pending:
            now = clock()
            if pending[0].arrival > now:
                time.sleep(min(pending[0].arrival - now, 0.0005))
                continue
            round_, mem = [], 0
            while pending and pending[0].arrival <= now:
                f = self.footprint(pending[0])
                if round_ and (mem + f > self.budget or not batched):
                    break
                round_.append(pending.popleft())
                mem += f
            self.run_round(round_, batched, clock)
        return stream


def summarize(stream: List[Request]) -> dict:
    lat = np.array([r.latency for r in stream]) * 1e3
    span = max(r.done for r in stream) - min(r.arrival for r in stream)
    return {"requests": len(stream), "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "qps": len(stream) / span if span > 0 else float("inf")}


def fresh(stream: List[Request]) -> List[Request]:
    """The same requests with no results (to replay a stream)."""
    return [Request(r.rid, r.shape, r.r0, r.r1, r.arrival) for r in stream]
