"""The allocation-free fused-aggregate probe (qdb_gpu_join_probe_fused_noalloc)
against the probe it varies (jr_dispatch Probe + qdb_gpu_join_fused_results),
through the C-ABI: two calls in a row on one join, the second with a smaller
batch, each equal to a fresh join's one probe of the same batch; no device
allocation (the pool's counters do not move) during the calls."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fresh_probe(W, spec, jd, batch):
    """(fused results, pair count) of ONE probe of `batch` by the existing
    entry points on a fresh join over the same build."""
    from qumirdb_amd.host import FilterKernel, RowJoinKernels, TRowJoinProcessor
    k = RowJoinKernels(spec)
    p = TRowJoinProcessor(k, build_is_left=False)
    p.append(jd.build)
    p.build()
    pairs = p.probe_pairs(batch)
    res = p.fused_results()
    p.close()
    return res, len(pairs)


@pytest.mark.parametrize("table_mode", [0, 1])
def test_noalloc_probe_equals_probe(gpu, table_mode):
    from qumirdb_amd import abi, workloads as W
    from qumirdb_amd.host import FilterKernel, JoinSpec, RowJoinKernels, TRowJoinProcessor
    lib = gpu.lib()
    print(f"no-alloc probe, table mode {table_mode}", flush=True)
    jd = W.JoinDevice(300_000, 3_000_000)
    sel = FilterKernel(W.BUILD_TYPES, *W.join_build_predicate())(jd.build)
    jd.build.set_selection_buffer(sel)
    spec = JoinSpec(build_types=W.BUILD_TYPES, probe_types=W.PROBE_TYPES, build_keys=[W.BUILD_KEY],
                    probe_keys=[W.PROBE_FK], join_type=abi.JOIN_INNER, probe_is_left=True,
                    fused=[(abi.AGG_COUNT_STAR, -1), (abi.AGG_SUM, W.PROBE_VALUE), (abi.AGG_MIN, W.PROBE_VALUE),
                           (abi.AGG_MAX, W.PROBE_VALUE)],
                    table_mode=table_mode)
    big, small = jd.probe, jd.probe.view(0, 700_001)
    want = [_fresh_probe(W, spec, jd, b) for b in (big, small)]

    k = RowJoinKernels(spec)
    p = TRowJoinProcessor(k, build_is_left=False)
    p.append(jd.build)
    p.build()
    arena = lib.qdb_gpu_join_processor_arena(p.p)
    stats0 = (C.c_int64 * 5)()
    lib.qdb_gpu_pool_stats(stats0, 5)
    for (wres, wcount), batch in zip(want, (big, small)):
        out = (abi.QdbFusedResult * 4)()
        seen = C.c_int64(-1)
        rc = lib.qdb_gpu_join_probe_fused_noalloc(arena, batch.ptr, out, 4, C.byref(seen))
        assert rc == 0, gpu.last_error()
        assert seen.value == wcount
        got = [{"value": (int(r.Hi) << 64) + (int(r.Lo) & ((1 << 64) - 1)), "count": int(r.Count), "valid": bool(r.Valid)}
               for r in out]
        assert got == wres
    stats1 = (C.c_int64 * 5)()
    lib.qdb_gpu_pool_stats(stats1, 5)
    # No driver allocation and no block handed out by the pool.
    assert stats1[0] == stats0[0] and stats1[4] == stats0[4], (list(stats0), list(stats1))
    # The probe's accumulated results (qdb_gpu_join_fused_results) are untouched.
    assert all(r["count"] == 0 for r in p.fused_results())
    p.close()


def test_noalloc_probe_rejects_join_without_fused(gpu):
    from qumirdb_amd import workloads as W
    from qumirdb_amd.host import RowJoinKernels, TRowJoinProcessor
    jd = W.JoinDevice(10_000, 10_000)
    p = TRowJoinProcessor(RowJoinKernels(W.join_spec(fused=False)), build_is_left=False)
    p.append(jd.build)
    p.build()
    out = (gpu.abi.QdbFusedResult * 4)()
    seen = C.c_int64(0)
    assert gpu.lib().qdb_gpu_join_probe_fused_noalloc(gpu.lib().qdb_gpu_join_processor_arena(p.p), jd.probe.ptr, out, 4,
                                                      C.byref(seen)) == -1
    assert "no fused aggregate" in gpu.last_error()
    p.close()
