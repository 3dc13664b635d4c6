"""Writes tests/golden/benchlog_reference_db.json: what the reference's own
benchmark-log reader makes of a log written by ompl_amd/benchlog.py.

Run in the build container only (it reads the reference tree):
    python tests/golden/make_benchlog_golden.py

The reader is readBenchmarkLog (py-bindings/ompl/_benchmark_statistics.py:110
of the reference), a standard-library-only script; it is imported by path,
run on the log below, and every row of the SQLite database it builds is
stored with the log text.  tests/test_benchlog.py then checks, without the
reference, that benchlog.py still writes this exact text and that the parsed
values are the reader's.
"""
from __future__ import annotations

import datetime as dt
import importlib.util
import json
import sqlite3
import sys
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
READER = Path("/root/reference/py-bindings/ompl/_benchmark_statistics.py")
OUT = HERE / "benchlog_reference_db.json"


def fixed_log() -> str:
    """A log with both planners of the served-stream comparison and every run
    property kind benchlog.py writes (fixed host, date and values)."""
    sys.path.insert(0, str(ROOT))
    from ompl_amd import benchlog

    runs = [
        {"time REAL": 0.005, "latency REAL": 0.006, "queued time REAL": 0.001, "status ENUM": 6,
         "solved BOOLEAN": True, "graph states INTEGER": 30, "iterations INTEGER": 12,
         "solution segments INTEGER": 2, "solution length REAL": 1.3228756555322954},
        {"time REAL": 0.08, "latency REAL": 0.0805, "queued time REAL": 0.0005, "status ENUM": 4,
         "solved BOOLEAN": False, "graph states INTEGER": 900, "iterations INTEGER": 2000},
        {"time REAL": 0.002, "latency REAL": 0.0021, "queued time REAL": 0.0001, "status ENUM": 8,
         "solved BOOLEAN": False, "graph states INTEGER": 20, "iterations INTEGER": 7},
    ]
    return benchlog.to_string(
        "served_vs_plain",
        [{"name": "geometric_RRTConnect_served", "common": {"range": 3.3, "longest valid segment": 0.01},
          "runs": runs},
         {"name": "geometric_RRTConnect_plain", "common": {}, "runs": runs[:2]}],
        seed=8, max_time=10.0, max_mem=4096.0, total_duration=1.25, host="mi355x-host",
        start=dt.datetime(2026, 12, 3, 10, 0, 0, 123456), setup_info="arm: 7-DOF sphere model\n",
        cpu_info="synthetic\n", properties={"qps REAL": 200.0, "requests INTEGER": 3, "environments INTEGER": 1})


def main() -> int:
    spec = importlib.util.spec_from_file_location("ompl_benchmark_statistics", READER)
    reader = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(reader)
    text = fixed_log()
    with tempfile.TemporaryDirectory() as d:
        log = Path(d) / "served.log"
        log.write_text(text)
        db = Path(d) / "bench.db"
        reader.readBenchmarkLog(str(db), [str(log)], False)
        conn = sqlite3.connect(str(db))
        tables = {}
        for (name,) in conn.execute("SELECT name FROM sqlite_master WHERE type = 'table' ORDER BY name"):
            if name.startswith("sqlite_"):
                continue
            cur = conn.execute(f"SELECT * FROM {name}")
            cols = [c[0] for c in cur.description]
            tables[name] = {"columns": cols, "rows": [list(r) for r in cur.fetchall()]}
    OUT.write_text(json.dumps({"reader": "py-bindings/ompl/_benchmark_statistics.py:110 readBenchmarkLog",
                               "log": text, "tables": tables}, indent=1, sort_keys=True) + "\n")
    counts = ", ".join("%s (%d)" % (k, len(v["rows"])) for k, v in tables.items())
    print(f"wrote {OUT}: {counts}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
