"""Writes tests/golden/fixtures.json: input tables and expected results that
the reference's own regression suite holds for the aggregate and join path.

Run in a container where the reference is mounted (it reads, never imports,
/root/reference). Only data is copied: the table rows (test/regtest/cases/
*/tables/*.json) and the golden result rows (test/regtest/goldens/*/
*.result.json). Each test states the query it stands for in its own words.
"""
import json
import os

REF = "/root/reference/test/regtest"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures.json")

TABLES = {
    "decimal_parquet.wide": "cases/decimal_parquet/tables/wide.json",
    "decimal_parquet.wide_dim": "cases/decimal_parquet/tables/wide_dim.json",
    "decimal_result.t": "cases/decimal_result/tables/t.json",
    "decimal_result.u": "cases/decimal_result/tables/u.json",
    "aggregate_row_layout.t": "cases/aggregate_row_layout/tables/t.json",
    "aggregate_empty.empty": "cases/aggregate_empty/tables/empty.json",
    "aggregate_empty.values": "cases/aggregate_empty/tables/values.json",
    "aggregate_empty.nulls": "cases/aggregate_empty/tables/nulls.json",
    "outer_join.c": "cases/outer_join/tables/c.json",
    "outer_join.o": "cases/outer_join/tables/o.json",
    "decimal_parquet.layout": "cases/decimal_parquet/tables/layout.json",
}

# Data files of the reference's tests copied as they are (a Parquet file of
# the decimal scan cases: FIXED_LEN_BYTE_ARRAY decimals of 4, 7, 8 and 16
# bytes, 16 row groups).
DATA_FILES = {"decimals.parquet": "cases/decimal_parquet/tables/decimals.parquet"}

GOLDENS = [
    "decimal_parquet/aggregates", "decimal_parquet/group_by", "decimal_parquet/casts", "decimal_parquet/join",
    "decimal_result/multiply_sum", "decimal_result/storage_sum", "decimal_result/multiply",
    "decimal_result/storage_boundary",
    "aggregate_row_layout/checksum", "aggregate_row_layout/estimated_groups",
    "aggregate_empty/empty_mixed", "aggregate_empty/nonempty_mixed", "aggregate_empty/all_null_mixed",
    "aggregate_empty/grouped_empty",
    "outer_join/duplicates", "outer_join/null_keys", "outer_join/all_matched", "outer_join/none_matched",
    "outer_join/right_join",
    "decimal_parquet/order_limit", "decimal_parquet/order_wide", "decimal_parquet/filter_bounds",
    "decimal_parquet/layouts_agree",
]

# Goldens whose expected outcome is a runtime error (the error text).
ERRORS = ["decimal_result/storage_overflow", "decimal_result/storage_unsigned"]


def main():
    out = {"tables": {}, "goldens": {}, "source": "QumirLabs/qumirdb test/regtest (data only)"}
    for name, rel in TABLES.items():
        with open(os.path.join(REF, rel)) as f:
            t = json.load(f)
        out["tables"][name] = {"schema": t["schema"], "rows": t["rows"], "from": rel}
    for g in GOLDENS:
        rel = f"goldens/{g}.result.json"
        with open(os.path.join(REF, rel)) as f:
            r = json.load(f)
        out["goldens"][g] = {"schema": r["schema"], "rows": r["rows"], "from": rel}
    out["errors"] = {}
    for g in ERRORS:
        rel = f"goldens/{g}.error"
        with open(os.path.join(REF, rel)) as f:
            out["errors"][g] = {"text": f.read().strip(), "from": rel}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    here = os.path.dirname(OUT)
    for name, rel in DATA_FILES.items():
        with open(os.path.join(REF, rel), "rb") as f:
            data = f.read()
        with open(os.path.join(here, name), "wb") as f:
            f.write(data)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
