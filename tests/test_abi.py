"""The C-ABI library loads and exports every entry point include/qdb_gpu.h
declares, with struct layouts that match the bindings. No compute call is
made (these run without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "qumirdb_amd", "libqdb_gpu.so")
HEADER = os.path.join(ROOT, "include", "qdb_gpu.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qdb_gpu_\w+)\s*\(", text)))


def test_library_exists():
    assert os.path.exists(LIB), "build the library first (__graft_entry__.build())"


def test_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [f for f in header_functions() if not hasattr(lib, f)]
    assert not missing, f"not exported: {missing}"
    assert len(header_functions()) >= 40


def test_exports_are_extern_c():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    names = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for f in header_functions():
        assert f in names, f"{f} is not an unmangled dynamic symbol"


def test_binding_covers_header():
    from qumirdb_amd import abi
    assert sorted(abi.EXPORTS) == header_functions()


def test_struct_sizes_match_bindings():
    from qumirdb_amd import abi
    lib = abi.declare(ctypes.CDLL(LIB))
    n = len(abi.STRUCTS)
    out = (ctypes.c_int64 * n)()
    count = lib.qdb_gpu_struct_sizes(out, n)
    assert count == n
    for s, size in zip(abi.STRUCTS, out):
        assert ctypes.sizeof(s) == size, f"{s.__name__}: binding {ctypes.sizeof(s)} vs library {size}"


def test_reference_blob_sizes():
    # qdb/kernel/compiler.h:603-616 and qdb/io/io.h (TColumn 48, TRowSet 64).
    from qumirdb_amd import abi
    assert (abi.HASH_TABLE_SIZE, abi.PAIR_BUFFER_SIZE, abi.ROW_ARENA_SIZE, abi.ROW_TABLE_SIZE, abi.ROW_SCRATCH_SIZE) == (304, 24, 56, 48, 64)
    assert ctypes.sizeof(abi.TColumn) == 48
    assert ctypes.sizeof(abi.TRowSet) == 64
    assert ctypes.sizeof(abi.QdbPairBuffer) == 24


def test_package_import_fails_loudly_without_library(tmp_path):
    # A copy of the package without its .so must refuse to import.
    import shutil
    pkg = tmp_path / "qumirdb_amd"
    shutil.copytree(os.path.join(ROOT, "qumirdb_amd"), pkg, ignore=shutil.ignore_patterns("*.so", "_build", "__pycache__"))
    code = "import sys; sys.path.insert(0, %r); import qumirdb_amd" % str(tmp_path)
    proc = subprocess.run(["python3", "-c", code], capture_output=True, text=True)
    assert proc.returncode != 0 and "no CPU fallback" in proc.stderr


def test_autotune_cache_file_round_trip(tmp_path):
    """The tuner's cache file: valid lines load, invalid ones (unknown LDS
    budget, fan-out out of range, malformed) are skipped, and a saved cache
    loads back to the same entries. No GPU call."""
    from qumirdb_amd import abi
    lib = abi.declare(ctypes.CDLL(LIB))
    lib.qdb_gpu_autotune_clear()
    src = tmp_path / "tune.txt"
    src.write_text("# kw card rows path lds grid shift blocks copies ms\n"
                   "1 14 18 0 64 2 0 0 8 1.5\n"
                   "1 21 24 1 0 0 -1 512 1 4.25\n"
                   "2 10 12 0 48 2 0 0 1 1.0\n"     # 48 KiB is not a candidate
                   "1 22 25 1 0 0 5 0 1 1.0\n"      # fan-out shift out of range
                   "1 23 25 1 0 0 0 4096 1 1.0\n"   # more partition workgroups than 2048
                   "1 24 25 0 64 2 0 0 64 1.0\n"    # more LDS copies than 16
                   "garbage\n")
    assert lib.qdb_gpu_autotune_load(str(src).encode()) == 2
    n = lib.qdb_gpu_autotune_entries(None, 0)
    assert n == 2
    ents = (abi.QdbTuneEntry * n)()
    lib.qdb_gpu_autotune_entries(ents, n)
    got = sorted((e.KeyWords, e.CardBucket, e.RowsBucket, e.Path, e.LdsKiB, e.GridPerCu, e.FanoutShift, e.PartBlocks, e.LdsCopies,
                  e.Ms) for e in ents)
    assert got == [(1, 14, 18, 0, 64, 2, 0, 0, 8, 1.5), (1, 21, 24, 1, 0, 0, -1, 512, 1, 4.25)]
    out = tmp_path / "saved.txt"
    assert lib.qdb_gpu_autotune_save(str(out).encode()) == 0
    lib.qdb_gpu_autotune_clear()
    assert lib.qdb_gpu_autotune_entries(None, 0) == 0
    assert lib.qdb_gpu_autotune_load(str(out).encode()) == 2
    ents2 = (abi.QdbTuneEntry * 2)()
    lib.qdb_gpu_autotune_entries(ents2, 2)
    assert sorted((e.KeyWords, e.CardBucket, e.RowsBucket, e.Path, e.LdsKiB, e.GridPerCu, e.FanoutShift, e.PartBlocks,
                   e.LdsCopies, e.Ms) for e in ents2) == got
    lib.qdb_gpu_autotune_clear()
