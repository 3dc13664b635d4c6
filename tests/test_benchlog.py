"""The benchmark log writer produces the layout ompl_benchmark_statistics.py
reads (tools/benchmark/src/Benchmark.cpp:306-431): parsed back here with the
same line rules (token positions, "<<<|" blocks, "name TYPE" properties,
"value; " run lines)."""
import io

from ompl_amd import benchlog


class Run:
    def __init__(self, status, it, nodes, q, r, lat):
        self.status, self.iterations, self.tree_nodes = status, it, nodes
        self.queued_ms, self.run_ms, self.latency_ms, self.path_len = q, r, lat, 0


def parse(text):
    f = io.StringIO(text)
    rd = f.readline
    tok = lambda: rd().split()
    assert tok()[:2] == ["OMPL", "version"]
    name = rd().split()[-1]
    nprops = int(tok()[0])
    props = dict(rd().strip().split(" = ") for _ in range(nprops))
    assert rd().startswith("Running on")
    assert rd().startswith("Starting at")
    for _ in range(2):
        assert rd().startswith("<<<|")
        while not rd().startswith("|>>>"):
            pass
    seed = tok()[0]
    assert tok()[-3:] == ["seconds", "per", "run"] and tok()[-3:] == ["MB", "per", "run"]
    assert tok()[-3:] == ["runs", "per", "planner"]
    assert tok()[-3:] == ["collect", "the", "data"]
    nenum = int(tok()[0])
    enums = [rd()[:-1].split("|") for _ in range(nenum)]
    planners = {}
    for _ in range(int(tok()[0])):
        pname = rd()[:-1]
        common = [rd() for _ in range(int(tok()[0]))]
        names = ["_".join(rd().split()[:-1]) for _ in range(int(tok()[0]))]
        runs = []
        for _ in range(int(tok()[0])):
            vals = rd()[:-1].split("; ")[:-1]
            runs.append(dict(zip(names, vals)))
        assert rd() == ".\n"
        planners[pname] = (common, runs)
    return name, props, seed, enums, planners


def test_log_round_trip():
    from ompl_amd import serve

    res = [Run(serve.SOLVED, 12, 30, 1.0, 5.0, 6.0), Run(serve.FAILED, 2000, 900, 0.5, 80.0, 80.5),
           Run(serve.DEADLINE, 7, 20, 0.1, 2.0, 2.1)]
    paths = [[[0.0] * 7, [0.3] * 7, [0.5] * 7], [], []]
    runs = benchlog.plan_runs(res, paths)
    text = benchlog.to_string("served_vs_plain", [
        {"name": "geometric_RRTConnect_served", "common": {"range": 3.3}, "runs": runs},
        {"name": "geometric_RRTConnect_plain", "common": {}, "runs": runs[:1]}],
        seed=8, max_time=10.0, properties={"qps INTEGER": 200})
    name, props, seed, enums, planners = parse(text)
    assert name == "served_vs_plain" and props == {"qps INTEGER": "200"} and seed == "8"
    assert enums[0][0] == "status" and enums[0][7] == "Exact solution" and len(enums[0]) == 11
    common, r = planners["geometric_RRTConnect_served"]
    assert common == ["range = 3.3\n"]
    assert [x["status"] for x in r] == ["6", "4", "4"]
    assert r[0]["solved"] == "1" and r[1]["solved"] == "0"
    assert r[0]["solution_segments"] == "2" and r[1]["solution_segments"] == ""
    assert abs(float(r[0]["solution_length"]) - (0.3 + 0.2) * 7 ** 0.5) < 1e-12
    assert float(r[1]["time"]) == 80.0 / 1e3 and float(r[2]["latency"]) == 2.1 / 1e3
    assert len(planners["geometric_RRTConnect_plain"][1]) == 1


def test_reference_reader_fixture():
    """benchlog.py still writes the exact log that the reference's own reader
    (readBenchmarkLog, py-bindings/ompl/_benchmark_statistics.py:110) was run
    on to make tests/golden/benchlog_reference_db.json, and the reader's
    database holds the values written: experiment header and properties,
    planner settings, the status enum, and every run property."""
    import importlib.util
    import json
    from pathlib import Path

    golden = Path(__file__).resolve().parent / "golden"
    spec = importlib.util.spec_from_file_location("make_benchlog_golden", golden / "make_benchlog_golden.py")
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    fix = json.loads((golden / "benchlog_reference_db.json").read_text())
    text = mk.fixed_log()
    assert text == fix["log"]
    t = fix["tables"]
    exp = dict(zip(t["experiments"]["columns"], t["experiments"]["rows"][0]))
    assert exp["name"] == "served_vs_plain" and exp["version"] == "OMPL 2.0.2" and exp["seed"] == "8"
    assert exp["timelimit"] == 10.0 and exp["memorylimit"] == 4096.0 and exp["runcount"] == 3
    assert exp["totaltime"] == 1.25 and exp["qps"] == 200.0 and exp["requests"] == 3
    assert [r[1] for r in t["plannerConfigs"]["rows"]] == ["geometric_RRTConnect_served", "geometric_RRTConnect_plain"]
    assert [r[2] for r in t["enums"]["rows"]] == benchlog.STATUS_NAMES
    cols = t["runs"]["columns"]
    runs = [dict(zip(cols, r)) for r in t["runs"]["rows"]]
    _, _, _, _, planners = parse(text)
    mine = planners["geometric_RRTConnect_served"][1] + planners["geometric_RRTConnect_plain"][1]
    assert len(runs) == len(mine) == 5
    for ref, got in zip(runs, mine):
        for k, v in got.items():
            if v == "":
                assert ref[k] is None, k
            else:
                assert float(ref[k]) == float(v), (k, ref[k], v)
