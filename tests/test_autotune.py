"""The aggregate autotuner: every configuration it can pick gives the oracle's
results (forced one at a time, on small and large tables, one batch and two
batches into one table), and tuning caches one configuration per shape
bucket that later updates of that bucket reuse."""
import ctypes as C
import os

import numpy as np
import pytest

import cases
from cases import Batch, GpuBackend, OracleBackend, random_agg_case
from helpers import result_rows, sort_key

pytestmark = pytest.mark.gpu

# (path, lds_kib, grid_per_cu, fanout_shift, part_blocks, lds_copies): every candidate of the tuner.
CONFIGS = [(0, 32, 4, 0, 0, 1), (0, 64, 2, 0, 0, 1), (0, 128, 1, 0, 0, 1), (0, 64, 2, 0, 0, 8)] + \
    [(1, 0, 0, f, b, 1) for b in (2048, 512) for f in (-1, 0, 1)]


@pytest.fixture(scope="module")
def gpu_backend():
    return GpuBackend()


def staged_case(case):
    """Aggregates the bucketed path stages (no i128 MIN/MAX)."""
    keep = [(f, c) for f, c in case.aggs if not (f in (cases.AGG_MIN, cases.AGG_MAX) and case.types[c].width == 16)]
    return cases.AggCase(case.types, case.keys, keep)


@pytest.mark.parametrize("cfg", CONFIGS, ids=lambda c: "p%d_l%d_g%d_f%d_b%d_c%d" % c)
@pytest.mark.parametrize("shape", ["small_table", "large_table"])
def test_forced_config_equals_oracle(gpu_backend, monkeypatch, cfg, shape):
    monkeypatch.setenv("QDB_GPU_TUNE_FORCE", ",".join(str(v) for v in cfg))
    if shape == "small_table":
        # 200K rows over a few thousand groups into a 4096-slot table, two batches.
        case, b = random_agg_case(41, 200_000)
        capacity = 1 << 12
    else:
        # 1.1M rows into a 2^21-slot table (made lazily), two batches.
        case, b = random_agg_case(42, 1_100_000)
        capacity = 1 << 21
    case = staged_case(case)
    got = sorted(result_rows(gpu_backend.aggregate(case, [b, b], capacity=capacity)), key=sort_key)
    want = sorted(result_rows(OracleBackend(fsum_mode=2).aggregate(case, [b, b])), key=sort_key)
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert g == w, f"{cfg} {shape}: {g} vs {w}"


def test_tuner_caches_one_config_per_shape(gpu_backend, monkeypatch):
    monkeypatch.delenv("QDB_GPU_TUNE_FORCE", raising=False)
    lib = gpu_backend.q.lib()
    abi = gpu_backend.q.abi
    lib.qdb_gpu_autotune_clear()
    was = lib.qdb_gpu_autotune_enable(1)
    try:
        case, b = random_agg_case(43, 300_000)
        case = staged_case(case)
        want = sorted(result_rows(OracleBackend(fsum_mode=2).aggregate(case, [b])), key=sort_key)
        got = sorted(result_rows(gpu_backend.aggregate(case, [b], capacity=1 << 14)), key=sort_key)
        assert got == want
        n = lib.qdb_gpu_autotune_entries(None, 0)
        assert n == 1
        ents = (abi.QdbTuneEntry * n)()
        assert lib.qdb_gpu_autotune_entries(ents, n) == n
        e = ents[0]
        assert e.CardBucket == 14 and e.RowsBucket == 18 and e.Ms > 0
        assert (e.Path, e.LdsKiB, e.GridPerCu, e.FanoutShift, e.PartBlocks, e.LdsCopies) in CONFIGS
        # The same shape again: no new entry, same results.
        got2 = sorted(result_rows(gpu_backend.aggregate(case, [b], capacity=1 << 14)), key=sort_key)
        assert got2 == want and lib.qdb_gpu_autotune_entries(None, 0) == 1
        # Save, clear, load.
        path = os.path.join(os.environ.get("TMPDIR", "/tmp"), f"qdb_tune_{os.getpid()}.txt")
        assert lib.qdb_gpu_autotune_save(path.encode()) == 0
        lib.qdb_gpu_autotune_clear()
        assert lib.qdb_gpu_autotune_entries(None, 0) == 0
        assert lib.qdb_gpu_autotune_load(path.encode()) == 1
        ents2 = (abi.QdbTuneEntry * 1)()
        lib.qdb_gpu_autotune_entries(ents2, 1)
        assert (ents2[0].Path, ents2[0].LdsKiB, ents2[0].GridPerCu, ents2[0].FanoutShift, ents2[0].PartBlocks,
                ents2[0].LdsCopies) == (e.Path, e.LdsKiB, e.GridPerCu, e.FanoutShift, e.PartBlocks, e.LdsCopies)
        os.unlink(path)
    finally:
        lib.qdb_gpu_autotune_enable(was)
        lib.qdb_gpu_autotune_clear()


@pytest.mark.parametrize("mixed", [False, True], ids=["few_groups", "mixed"])
def test_partial_records_equal_oracle(gpu_backend, monkeypatch, mixed):
    """Tables of few groups under a large grid: workgroups write their groups
    as partial records and a second launch combines them. Mixed: the rows of
    half the workgroups (rows are dealt to workgroups in 256-row chunks,
    round robin) carry a few keys, the other half's thousands, so some
    workgroups write records and others update the 64-slot table themselves,
    which then grows and replays what did not fit (rows and combined records)."""
    import ctypes as C
    lib = gpu_backend.q.lib()
    limits = (C.c_int64 * 2)()
    lib.qdb_gpu_device_limits(limits, 2)
    grid = int(limits[0]) * 4
    monkeypatch.setenv("QDB_GPU_TUNE_FORCE", "0,32,4,0,0,1")
    n = grid * 256 * 2 + 1234
    case, b = random_agg_case(44, n)
    case = staged_case(case)
    rng = np.random.default_rng(45)
    chunk_wg = (np.arange(n) // 256) % grid
    k2 = rng.integers(0, 4, n).astype(np.int64)
    if mixed:
        many = chunk_wg >= grid // 2
        k2[many] = rng.integers(0, 3000, int(many.sum()))
    b.columns[1] = k2
    got = sorted(result_rows(gpu_backend.aggregate(case, [b, b], capacity=64)), key=sort_key)
    want = sorted(result_rows(OracleBackend(fsum_mode=2).aggregate(case, [b, b])), key=sort_key)
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert g == w, f"{g} vs {w}"
