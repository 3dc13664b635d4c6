"""Writer of OMPL benchmark log files for planning runs made with ompl_amd.

The format is the one ompl::tools::Benchmark::saveResultsToStream writes
(tools/benchmark/src/Benchmark.cpp:306-431) and ompl_benchmark_statistics.py
reads into the Planner Arena database: a header (version, experiment name,
experiment properties, host, start time, setup and CPU blocks, seed, limits,
run count, total duration), the status enum, then per planner its common
properties, the sorted names of its run properties ("name TYPE"), one line per
run with "value; " per property (empty when a run lacks it) and a final ".".

Run properties use the reference's names where they exist ("time REAL",
"solved BOOLEAN", "status ENUM", "valid segment fraction REAL", "graph states
INTEGER", "solution segments INTEGER", "solution length REAL"); served-path
extras are "latency REAL" and "queued time REAL" (seconds).
"""
from __future__ import annotations

import datetime as _dt
import io
import math
import platform
from typing import Iterable, TextIO

__all__ = ["STATUS_NAMES", "status_of_plan", "write_log", "plan_runs"]

OMPL_VERSION = "2.0.2"

# PlannerStatus::StatusType order (base/PlannerStatus.h:51-75) and asString()
# (base/src/PlannerStatus.cpp:39-60; ABORT and INFEASIBLE print as the default)
STATUS_NAMES = ["Unknown status", "Invalid start", "Invalid goal", "Unrecognized goal type", "Timeout",
                "Approximate solution", "Exact solution", "Crash", "Unknown status", "Unknown status"]
TIMEOUT, EXACT_SOLUTION, ABORT = 4, 6, 8


def status_of_plan(status: int) -> int:
    """ompl_amd plan status -> PlannerStatus::StatusType value."""
    from . import serve

    return {serve.SOLVED: EXACT_SOLUTION, serve.FAILED: TIMEOUT, serve.DEADLINE: TIMEOUT,
            serve.CANCELLED: ABORT, serve.REJECTED: ABORT, serve.TREE_FULL: ABORT}.get(status, 0)


def _fmt(v) -> str:
    if isinstance(v, bool):
        return "1" if v else "0"
    if isinstance(v, float):
        return repr(v) if math.isfinite(v) else ("inf" if v > 0 else "nan")
    return str(v)


def write_log(out: TextIO | str, experiment: str, planners: list[dict], *, seed: int = 0, max_time: float = 0.0,
              max_mem: float = 0.0, total_duration: float = 0.0, host: str | None = None,
              start: _dt.datetime | None = None, setup_info: str = "", cpu_info: str = "",
              properties: dict | None = None) -> None:
    """planners: [{"name": str, "common": {name: value}, "runs": [{"prop TYPE": value, ...}, ...]}].
    properties: experiment properties {"name TYPE": value}."""
    if isinstance(out, str):
        with open(out, "w") as f:
            return write_log(f, experiment, planners, seed=seed, max_time=max_time, max_mem=max_mem,
                             total_duration=total_duration, host=host, start=start, setup_info=setup_info,
                             cpu_info=cpu_info, properties=properties)
    props = properties or {}
    runs_per = max((len(p["runs"]) for p in planners), default=0)
    w = out.write
    w(f"OMPL version {OMPL_VERSION}\n")
    w(f"Experiment {experiment or 'NO_NAME'}\n")
    w(f"{len(props)} experiment properties\n")
    for k, v in props.items():
        w(f"{k} = {_fmt(v)}\n")
    w(f"Running on {host or platform.node() or 'UNKNOWN'}\n")
    w(f"Starting at {(start or _dt.datetime.now()).strftime('%Y-%b-%d %H:%M:%S.%f')}\n")
    w(f"<<<|\n{setup_info}|>>>\n")
    w(f"<<<|\n{cpu_info}|>>>\n")
    w(f"{seed} is the random seed\n")
    w(f"{_fmt(float(max_time))} seconds per run\n")
    w(f"{_fmt(float(max_mem))} MB per run\n")
    w(f"{runs_per} runs per planner\n")
    w(f"{_fmt(float(total_duration))} seconds spent to collect the data\n")
    w("1 enum type\n")
    w("status|" + "|".join(STATUS_NAMES) + "\n")
    w(f"{len(planners)} planners\n")
    for p in planners:
        w(p["name"] + "\n")
        common = p.get("common", {})
        w(f"{len(common)} common properties\n")
        for k in sorted(common):
            w(f"{k} = {_fmt(common[k])}\n")
        names = sorted({k for r in p["runs"] for k in r})
        w(f"{len(names)} properties for each run\n")
        for k in names:
            w(k + "\n")
        w(f"{len(p['runs'])} runs\n")
        for r in p["runs"]:
            w("".join((_fmt(r[k]) if k in r else "") + "; " for k in names) + "\n")
        w(".\n")


def plan_runs(results: Iterable, paths: Iterable | None = None) -> list[dict]:
    """Run property dicts for plan results (ompl_amd.serve.PlanResult) and,
    optionally, their paths ([n, 7] arrays)."""
    import numpy as np

    runs = []
    paths = list(paths) if paths is not None else None
    for i, r in enumerate(results):
        run = {"time REAL": r.run_ms / 1e3, "latency REAL": r.latency_ms / 1e3, "queued time REAL": r.queued_ms / 1e3,
               "status ENUM": status_of_plan(r.status), "solved BOOLEAN": status_of_plan(r.status) == EXACT_SOLUTION,
               "graph states INTEGER": int(r.tree_nodes), "iterations INTEGER": int(r.iterations)}
        if paths is not None and len(paths[i]) > 0:
            p = np.asarray(paths[i], np.float64)
            run["solution segments INTEGER"] = max(len(p) - 1, 0)
            run["solution length REAL"] = float(np.sqrt(((p[1:] - p[:-1]) ** 2).sum(1)).sum()) if len(p) > 1 else 0.0
        runs.append(run)
    return runs


def to_string(*args, **kw) -> str:
    s = io.StringIO()
    write_log(s, *args, **kw)
    return s.getvalue()
