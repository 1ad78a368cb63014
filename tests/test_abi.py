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
