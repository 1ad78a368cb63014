"""Test helpers: the reference's golden fixtures as typed columns, and result
comparison."""
from __future__ import annotations

import json
import os
import re
from decimal import Decimal

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = os.path.join(HERE, "golden", "fixtures.json")

# Value kinds of the C-ABI (include/qdb_gpu.h).
SIGNED, UNSIGNED, FLOAT, BOOL, STRING, DATE, DECIMAL = range(7)


class T:
    """A column type: kind, width, scale, nullable, precision."""

    def __init__(self, kind, width, scale=0, nullable=False, precision=0):
        self.kind, self.width, self.scale, self.nullable, self.precision = kind, width, scale, nullable, precision

    def __repr__(self):
        return f"T({self.kind},{self.width},s{self.scale},{'null' if self.nullable else 'nn'})"


def load_fixtures():
    with open(FIXTURES) as f:
        return json.load(f)


_FIELD = re.compile(r"\((\w+) (<named Nullable \[(.+?)\]>|<named DECIMAL \[(\d+) (\d+)\]>|\w+)\)")


def _leaf(text: str, nullable: bool) -> T:
    m = re.match(r"<named DECIMAL \[(\d+) (\d+)\]>", text)
    if m:
        p, s = int(m.group(1)), int(m.group(2))
        return T(DECIMAL, 8 if p <= 18 else 16, s, nullable, p)
    return {
        "i8": T(SIGNED, 1, nullable=nullable), "i16": T(SIGNED, 2, nullable=nullable),
        "i32": T(SIGNED, 4, nullable=nullable), "i64": T(SIGNED, 8, nullable=nullable),
        "u64": T(UNSIGNED, 8, nullable=nullable),
        "f64": T(FLOAT, 8, nullable=nullable), "string": T(STRING, 0, nullable=nullable),
        "bool": T(BOOL, 1, nullable=nullable),
    }[text]


def parse_schema(schema: str):
    """[(name, T)] of a '<struct (name type) ...>' schema string."""
    body = schema[len("<struct "):-1]
    out = []
    i = 0
    while i < len(body):
        if body[i] != "(":
            i += 1
            continue
        depth, j = 0, i
        while True:
            if body[j] == "(":
                depth += 1
            elif body[j] == ")":
                depth -= 1
                if depth == 0:
                    break
            j += 1
        field = body[i + 1:j]
        name, typ = field.split(" ", 1)
        if typ.startswith("<named Nullable ["):
            out.append((name, _leaf(typ[len("<named Nullable ["):-2], True)))
        else:
            out.append((name, _leaf(typ, False)))
        i = j + 1
    return out


def decimal_unscaled(text, scale: int) -> int:
    """The unscaled integer of a decimal literal at `scale` (exact, any
    number of digits)."""
    s = str(text).strip()
    neg = s.startswith("-")
    s = s.lstrip("+-")
    whole, _, frac = s.partition(".")
    if len(frac) > scale:
        assert int(frac[scale:] or "0") == 0, f"{text} has more than {scale} decimals"
        frac = frac[:scale]
    v = int(whole or "0") * 10 ** scale + int((frac + "0" * scale)[:scale] or "0")
    return -v if neg else v


def table_columns(table, drop_strings=True):
    """(names, types, values, valid) of a fixture table; decimals as unscaled
    ints, integers as ints, NULL as 0 with valid False."""
    fields = parse_schema(table["schema"])
    rows = table["rows"]
    names, types, values, valid = [], [], [], []
    for ci, (name, t) in enumerate(fields):
        if t.kind == STRING and drop_strings:
            continue
        vals, ok = [], []
        for r in rows:
            v = r[ci]
            ok.append(v is not None)
            if v is None:
                vals.append(0)
            elif t.kind == DECIMAL:
                vals.append(decimal_unscaled(v, t.scale))
            elif t.kind == FLOAT:
                vals.append(float(v))
            else:
                vals.append(int(v))
        names.append(name)
        types.append(t)
        values.append(vals)
        valid.append(np.array(ok, dtype=bool) if t.nullable else None)
    return names, types, values, valid


def np_column(t: T, vals):
    if t.width == 16:
        return vals
    if t.kind == FLOAT:
        return np.array(vals, dtype=np.float64)
    if t.kind == UNSIGNED:
        return np.array(vals, dtype={1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[t.width])
    return np.array(vals, dtype={1: np.int8, 2: np.int16, 4: np.int32, 8: np.int64}[t.width])


def golden_rows(golden):
    """Golden rows as tuples: ints for integer/decimal (unscaled), floats,
    None for NULL."""
    fields = parse_schema(golden["schema"])
    out = []
    for r in golden["rows"]:
        row = []
        for (name, t), v in zip(fields, r):
            if v is None:
                row.append(None)
            elif t.kind == DECIMAL:
                row.append(decimal_unscaled(v, t.scale))
            elif t.kind == FLOAT:
                row.append(float(v))
            elif t.kind == STRING:
                row.append(v)
            else:
                row.append(int(v))
        out.append(tuple(row))
    return out


def result_rows(columns):
    """Rows of (values, valid) columns as tuples with None for NULL."""
    if not columns:
        return []
    n = len(columns[0][0])
    rows = []
    for g in range(n):
        row = []
        for vals, valid in columns:
            if valid is not None and not valid[g]:
                row.append(None)
            else:
                v = vals[g]
                row.append(float(v) if isinstance(v, (float, np.floating)) else int(v))
        rows.append(tuple(row))
    return rows


def sort_key(row):
    return tuple((0, 0) if v is None else (1, v) for v in row)
