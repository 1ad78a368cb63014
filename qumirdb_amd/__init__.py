"""MI355X (gfx950) hash-aggregate / hash-join path for QumirDB.

The product is the HIP library libqdb_gpu.so (built in-tree by build.py); this
package loads it through its C-ABI (include/qdb_gpu.h) and offers a thin
host-side mirror of the reference's operator interface for the path. There is
no CPU fallback: importing the package fails when the library is missing.
"""
from __future__ import annotations

import ctypes
import os

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
# QDB_GPU_LIB selects another build of the library (same C-ABI).
LIB_PATH = os.environ.get("QDB_GPU_LIB") or os.path.join(_HERE, "libqdb_gpu.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"qumirdb_amd: {LIB_PATH} is missing; build it with `python qumirdb_amd/build.py` "
        "(or __graft_entry__.build()). The HIP library is required: there is no CPU fallback.")

_lib = abi.declare(ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL))


def lib() -> ctypes.CDLL:
    """The loaded HIP library."""
    return _lib


def last_error() -> str:
    return (_lib.qdb_gpu_last_error() or b"").decode()


def exec_last_error() -> str:
    return (_lib.qdb_gpu_exec_last_error() or b"").decode()


def check(rc: int, what: str) -> int:
    if rc < 0:
        raise RuntimeError(f"{what}: {last_error() or exec_last_error()}")
    return rc


from .host import (  # noqa: E402  (needs _lib)
    AggregateKernels, DeviceBuffer, DeviceRowSet, FilterKernel, Program, ProjectKernel, RowJoinKernels,
    TAggregateProcessor, TRowJoinProcessor, ValueType, decimal_add_type, decimal_mul_type, download_rowset,
    init_device,
)

__all__ = [
    "abi", "lib", "last_error", "check", "AggregateKernels", "DeviceBuffer", "DeviceRowSet", "FilterKernel",
    "Program", "ProjectKernel", "RowJoinKernels", "TAggregateProcessor", "TRowJoinProcessor", "ValueType",
    "decimal_add_type", "decimal_mul_type", "download_rowset", "init_device",
]
