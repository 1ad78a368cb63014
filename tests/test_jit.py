"""The update kernels generated per aggregate (compiled at run time with
hipRTC, as the reference's compiler generates a kernel per query,
qdb/kernel/compiler.cpp): for every aggregate the parity suite runs, and
the three benchmark aggregates, the generated source compiles for gfx950.
No GPU is needed: the source is generated and compiled on the host."""
import ctypes as C

import pytest

import cases
from cases import OracleBackend, random_agg_case


class CaptureBackend(OracleBackend):
    """The oracle, recording the aggregate specs it is asked to run."""

    def __init__(self):
        super().__init__(fsum_mode=2)
        self.specs = []

    def aggregate(self, case, batches, **kw):
        self.specs.append(case.to_c())
        return super().aggregate(case, batches)


def _parity_specs():
    cap = CaptureBackend()
    for name in sorted(cases.AGG_GOLDENS):
        cases.AGG_GOLDENS[name](cap)
    specs = [("golden", s) for s in cap.specs]
    specs.append(("random", random_agg_case(0, 16)[0].to_c()))
    return specs


def _workload_specs():
    from qumirdb_amd import workloads as W
    return [("lineitem_fused", W.lineitem_fused_spec().to_c()), ("lineitem_plain", W.lineitem_plain_spec().to_c()),
            ("groups", W.groups_spec().to_c())]


def _source(lib, q, spec):
    s = getattr(q.abi, type(spec).__name__).from_buffer_copy(spec)
    n = lib.qdb_gpu_aggregate_kernel_source(C.byref(s), None, 0)
    assert n > 0, q.last_error()
    buf = C.create_string_buffer(n + 1)
    assert lib.qdb_gpu_aggregate_kernel_source(C.byref(s), buf, n + 1) == n
    return buf.value


def test_generated_update_kernels_compile():
    import qumirdb_amd as q
    lib = q.lib()
    seen = set()
    for name, spec in _parity_specs() + _workload_specs():
        src = _source(lib, q, spec)
        if src in seen:
            continue
        seen.add(src)
        assert b"qdb_agg_update" in src and b"JitPolicy" in src
        err = C.create_string_buffer(1 << 16)
        rc = lib.qdb_gpu_jit_check(src, b"gfx950", err, len(err))
        assert rc == 0, f"{name}: {err.value.decode(errors='replace')[:3000]}"
    assert len(seen) >= 8


def test_generated_source_fixes_the_layout():
    """The generated policy carries the plan's state-word operations as
    constants (no descriptor is read at run time), one straight-line read
    per argument column in the row contribution and in the staging of the
    bucketed path, and the bucketed path's kernels."""
    import qumirdb_amd as q
    from qumirdb_amd import workloads as W
    src = _source(q.lib(), q, W.groups_spec().to_c()).decode()
    # COUNT(*), SUM(i64), MIN/MAX(f64), SUM(f64) = rows, sum, min, max, 4 limbs, side word
    assert "static constexpr int kNW = 9;" in src
    assert src.count("read_col_t<") == 4  # contribution and staging
    assert "A.words[" not in src and "eval_program" not in src
    for k in ("qdb_agg_update", "qdb_agg_batch", "qdb_agg_count", "qdb_agg_scatter", "qdb_agg_sorted", "qdb_agg_merge"):
        assert k in src, k
    assert "static constexpr int kStage = 3;" in src  # key, SUM(i64) argument, f64 argument


def _filter_specs():
    """Filter predicates of the join workload and a three-valued OR over
    nullable columns (NULL is not TRUE)."""
    import qumirdb_amd as q
    from qumirdb_amd import workloads as W
    from qumirdb_amd.host import schema_c
    out = []
    prog, res = W.join_build_predicate()
    spec = q.abi.QdbFilterSpec()
    spec.Input = schema_c(W.BUILD_TYPES)
    spec.Predicate = prog.build(res)
    out.append(("join build", spec))
    E = cases._abi()
    from helpers import SIGNED, T
    types = [T(SIGNED, 8, nullable=True), T(SIGNED, 8, nullable=True)]
    p = cases.Prog(types)
    zero = p.const(0, T(SIGNED, 8))
    a = p.cmp(E.CMP_GT, p.col(0), zero)
    b = p.cmp(E.CMP_GT, p.col(1), zero)
    r = p._emit(E.EXPR_OR, E.REG_BOOL, T(3, 1, nullable=True), a=a, b=b)
    spec2 = q.abi.QdbFilterSpec()
    spec2.Input = q.abi.QdbColumnSchema.from_buffer_copy(cases.schema_c(types))
    spec2.Predicate = q.abi.QdbExprProgram.from_buffer_copy(p.build(r))
    out.append(("three-valued or", spec2))
    return out


def test_generated_filter_kernels_compile():
    """The kernel generated for a filter's predicate: straight-line code,
    one kernel qdb_filter, compiled for gfx950 on the host."""
    import qumirdb_amd as q
    lib = q.lib()
    for name, spec in _filter_specs():
        n = lib.qdb_gpu_filter_kernel_source(C.byref(spec), None, 0)
        assert n > 0, q.last_error()
        buf = C.create_string_buffer(n + 1)
        assert lib.qdb_gpu_filter_kernel_source(C.byref(spec), buf, n + 1) == n
        src = buf.value
        assert b"qdb_filter" in src and b"eval_program" not in src
        err = C.create_string_buffer(1 << 16)
        rc = lib.qdb_gpu_jit_check(src, b"gfx950", err, len(err))
        assert rc == 0, f"{name}: {err.value.decode(errors='replace')[:3000]}"


class ProjectCapture(OracleBackend):
    """The oracle, recording the project specs it is asked to run."""

    def __init__(self):
        super().__init__()
        self.specs = []

    def project(self, types, outputs, batch):
        self.specs.append(_project_spec(types, outputs))
        return super().project(types, outputs, batch)


def _project_spec(types, outputs):
    abi = cases._abi()
    spec = abi.QdbProjectSpec()
    spec.Input = cases.schema_c(types)
    spec.NumOutputs = len(outputs)
    for k, (p, r, t) in enumerate(outputs):
        spec.OutputTypes[k] = cases.value_type_c(t)
        spec.Outputs[k] = p.build(r)
    return spec


def _project_specs():
    """The golden projections, the lineitem products, and outputs of every
    store form: i128 decimal, f64, f32, a BOOL bitmap, a 2-byte integer."""
    import qumirdb_amd as q
    from qumirdb_amd import workloads as W
    from qumirdb_amd.host import schema_c
    from helpers import DECIMAL, FLOAT, SIGNED, T
    cap = ProjectCapture()
    for name in sorted(cases.PROJECT_GOLDENS):
        cases.PROJECT_GOLDENS[name](cap)
    out = [("golden", q.abi.QdbProjectSpec.from_buffer_copy(s)) for s in cap.specs]
    spec = q.abi.QdbProjectSpec()
    spec.Input = schema_c(W.LINEITEM_TYPES)
    spec.NumOutputs = 2
    for k, (prog, res) in enumerate([W.lineitem_disc_price(), W.lineitem_charge()]):
        spec.OutputTypes[k] = prog.result_type(res).to_c()
        spec.Outputs[k] = prog.build(res)
    out.append(("lineitem", spec))
    E = cases._abi()
    types = [T(SIGNED, 8, nullable=True), T(DECIMAL, 8, 2, True, 15), T(FLOAT, 8, nullable=True), T(SIGNED, 2)]
    q1 = cases.Prog(types)
    x = q1.op(E.EXPR_MUL, q1.col(1), q1.const(100, T(DECIMAL, 8, 2, False, 3)), T(DECIMAL, 16, 4, True, 31))
    q2 = cases.Prog(types)
    y = q2.op(E.EXPR_ADD, q2.col(2), q2.const_f64(0.25), T(FLOAT, 8, nullable=True), cls=E.REG_F64)
    q3 = cases.Prog(types)
    z = q3.cmp(E.CMP_NE, q3.col(3), q3.const(0, T(SIGNED, 8)))
    q4 = cases.Prog(types)
    w = q4.col(3)
    outs = [(q1, x, T(DECIMAL, 16, 4, True, 31)), (q2, y, T(FLOAT, 8, nullable=True)), (q2, y, T(FLOAT, 4, nullable=True)),
            (q3, z, T(3, 1)), (q4, w, T(SIGNED, 2))]
    out.append(("store forms", q.abi.QdbProjectSpec.from_buffer_copy(_project_spec(types, outs))))
    return out


def test_generated_project_kernels_compile():
    """The kernel generated for a projection: each output's program as
    straight-line code with its store specialized to the output type, one
    kernel qdb_project, compiled for gfx950 on the host."""
    import qumirdb_amd as q
    lib = q.lib()
    for name, spec in _project_specs():
        n = lib.qdb_gpu_project_kernel_source(C.byref(spec), None, 0)
        assert n > 0, q.last_error()
        buf = C.create_string_buffer(n + 1)
        assert lib.qdb_gpu_project_kernel_source(C.byref(spec), buf, n + 1) == n
        src = buf.value
        assert b"qdb_project" in src and b"eval_program" not in src
        err = C.create_string_buffer(1 << 16)
        rc = lib.qdb_gpu_jit_check(src, b"gfx950", err, len(err))
        assert rc == 0, f"{name}: {err.value.decode(errors='replace')[:3000]}"
