"""Text summaries of rocprofv3 csv output, for the profile records.

    python -m ompl_amd.rocprof_summary stats DIR
        per-kernel calls and average / min / max duration (kernel_stats.csv, or
        the SQLite results database)
    python -m ompl_amd.rocprof_summary traffic BENCH_FETCH BENCH_WRITE COPY_FETCH COPY_WRITE COPY_BYTES
        per-kernel HBM bytes per dispatch from FETCH_SIZE / WRITE_SIZE runs, each
        scaled by what the counter reads for a plain device copy of COPY_BYTES
    python -m ompl_amd.rocprof_summary counters DIR [DIR ...] [--match TEXT ...]
        per-kernel average of every counter per dispatch, over one or more
        --pmc runs of the same program (kernels whose name holds a TEXT)
"""
from __future__ import annotations

import csv
import glob
import json
import sys
from collections import defaultdict


def _find(d: str, suffix: str) -> str:
    f = sorted(glob.glob(f"{d}/**/*{suffix}", recursive=True))
    if not f:
        raise FileNotFoundError(f"no *{suffix} under {d}")
    return f[0]


def kernel_stats_db(d: str) -> list[dict]:
    """The same statistics from a rocprofv3 SQLite results database (the
    default output format of ROCm 7.x: its `kernels` view, durations in ns)."""
    import sqlite3

    db = sqlite3.connect(_find(d, "_results.db"))
    rows = db.execute("SELECT name, COUNT(*), SUM(duration), AVG(duration), MIN(duration), MAX(duration) "
                      "FROM kernels GROUP BY name ORDER BY SUM(duration) DESC").fetchall()
    total = sum(r[2] for r in rows) or 1.0
    return [{"kernel": r[0], "calls": int(r[1]), "avg_us": r[3] / 1e3, "min_us": r[4] / 1e3, "max_us": r[5] / 1e3,
             "pct": 100.0 * r[2] / total} for r in rows]


def kernel_stats(d: str) -> list[dict]:
    if not glob.glob(f"{d}/**/*kernel_stats.csv", recursive=True):
        return kernel_stats_db(d)
    rows = list(csv.DictReader(open(_find(d, "kernel_stats.csv"))))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    return [{"kernel": r["Name"], "calls": int(r["Calls"]), "avg_us": float(r["AverageNs"]) / 1e3,
             "min_us": float(r["MinNs"]) / 1e3, "max_us": float(r["MaxNs"]) / 1e3,
             "pct": float(r["Percentage"])} for r in rows]


def counter_per_dispatch(d: str, counter: str) -> dict[str, float]:
    acc = defaultdict(list)
    for r in csv.DictReader(open(_find(d, "counter_collection.csv"))):
        if r["Counter_Name"] == counter:
            acc[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def counters_per_dispatch(dirs: list[str]) -> dict[str, dict[str, float]]:
    """{kernel: {counter: average value per dispatch}} over the --pmc runs in
    `dirs` (each counter is taken from the run that collected it)."""
    acc: dict[str, dict[str, list[float]]] = defaultdict(lambda: defaultdict(list))
    for d in dirs:
        # a dispatch's value of a counter is the sum of its rows (one row per
        # dimension instance when the counter is not reduced)
        per: dict[tuple, float] = defaultdict(float)
        for r in csv.DictReader(open(_find(d, "counter_collection.csv"))):
            key = (r["Kernel_Name"], r["Counter_Name"], r.get("Dispatch_Id", r.get("Correlation_Id", "")))
            per[key] += float(r["Counter_Value"])
        for (k, c, _), v in per.items():
            acc[k][c].append(v)
    return {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in acc.items()}


def traffic(bench_fetch: str, bench_write: str, copy_fetch: str, copy_write: str, copy_bytes: int) -> dict:
    cf = counter_per_dispatch(copy_fetch, "FETCH_SIZE")
    cw = counter_per_dispatch(copy_write, "WRITE_SIZE")
    ck = [k for k in cf if "copyBuffer" in k or "copy" in k.lower()]
    if not ck:
        raise ValueError("no copy kernel in the calibration runs")
    k = ck[0]
    f_scale = copy_bytes / (cf[k] * 1024.0)  # bytes per counted KiB
    w_scale = copy_bytes / (cw[k] * 1024.0)
    bf = counter_per_dispatch(bench_fetch, "FETCH_SIZE")
    bw = counter_per_dispatch(bench_write, "WRITE_SIZE")
    out = {"copy_kernel": k, "copy_bytes": copy_bytes, "fetch_scale": f_scale, "write_scale": w_scale, "kernels": {}}
    for name in sorted(set(bf) | set(bw), key=lambda n: -(bf.get(n, 0) + bw.get(n, 0))):
        rd = bf.get(name, 0.0) * 1024.0 * f_scale
        wr = bw.get(name, 0.0) * 1024.0 * w_scale
        out["kernels"][name] = {"read_bytes": rd, "write_bytes": wr, "bytes": rd + wr}
    return out


def main(argv: list[str]) -> int:
    if argv[0] == "stats":
        print(f"{'kernel':78s} {'calls':>6s} {'avg_us':>9s} {'min_us':>9s} {'max_us':>9s} {'%time':>6s}")
        for r in kernel_stats(argv[1]):
            print(f"{r['kernel'][:78]:78s} {r['calls']:6d} {r['avg_us']:9.2f} {r['min_us']:9.2f} {r['max_us']:9.2f} "
                  f"{r['pct']:6.2f}")
    elif argv[0] == "traffic":
        t = traffic(argv[1], argv[2], argv[3], argv[4], int(argv[5]))
        print(f"calibration: {t['copy_kernel']} of {t['copy_bytes']} B: FETCH_SIZE x {t['fetch_scale']:.3f}, "
              f"WRITE_SIZE x {t['write_scale']:.3f} (bytes per counted KiB / 1024)")
        print(f"{'kernel':78s} {'read_B':>12s} {'write_B':>12s} {'total_B':>12s}")
        for k, v in t["kernels"].items():
            print(f"{k[:78]:78s} {v['read_bytes']:12.0f} {v['write_bytes']:12.0f} {v['bytes']:12.0f}")
        print("json: " + json.dumps(t))
    elif argv[0] == "counters":
        dirs = [a for a in argv[1:] if not a.startswith("--")]
        match = []
        if "--match" in argv:
            i = argv.index("--match")
            dirs = [a for a in argv[1:i]]
            match = argv[i + 1:]
        t = counters_per_dispatch(dirs)
        names = sorted({c for cs in t.values() for c in cs})
        for k in sorted(t):
            if match and not any(m in k for m in match):
                continue
            print(k[:110])
            for c in names:
                if c in t[k]:
                    print(f"    {c:34s} {t[k][c]:16.1f}")
        print("json: " + json.dumps({k: v for k, v in t.items() if not match or any(m in k for m in match)}))
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
