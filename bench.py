"""Benchmark of the planning inner loop on one MI355X.

Workload (one "step" = one batched DiscreteMotionValidator pass): the
uneven-motion batch -- 7-DOF sphere arm, 256 obstacles (128 spheres r~U[0.05,0.15]
+ 128 axis-aligned boxes, half-extents U[0.05,0.2]; centres U[-1,1]^3 at least
0.2 m from the base), 16,384 edges with endpoints U[-pi,pi]^7, directions
uniform on S^6, lengths lognormal(mu=-1, sigma=1) clipped to [0.05, 5] rad,
0.01 rad resolution (5-500 states per edge), bisection order with first-collision
exit.  Inputs are resident in HBM before the timed region.

Reported (one JSON line):
  value            batched edge-checks/s (edges validated per second), the K
                   steps issued round-robin on --streams HIP streams (4, kept
                   only if their batches overlap), with the launch shape tuned
                   for those streams: independent batches, each fully checked;
                   the one- and two-stream figures are reported beside it
  plain            the same edges through the one-launch-per-call path
  roofline         dominant kernel: algorithmic FP32 work in the reference order
                   / its average duration (HIP events on its launch stream),
                   for the headline's shape and for the one-stream shape
  cpu_baseline     the CPU restatement on the host's 16-thread share, on a
                   bounded sample (one thread beside it)
  parity           batched results bit-exact vs the CPU restatement

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

N_EDGES = 16384
N_SPHERES, N_BOXES = 128, 128
ENV_SEED, EDGE_SEED = 1001, 4
PRM_SEED, PRM_SAMPLES = 7, 200_000
PEAK_FP32_TFLOPS = 157.3


def lvs_0p01():
    ext = 2 * math.pi * math.sqrt(7)
    return ext * (0.01 / ext)


# Threads of the CPU baselines: the host's CPU share of one GPU on the
# benchmark machines (16), or fewer when the host has fewer.
CPU_SHARE = 16
CPU_FLAGS = "C restatement, gcc -O3 -march=x86-64-v3 -ffp-contract=off"


def host_cpu():
    """The host CPU the baselines ran on: model and clock (lscpu / cpuinfo), nproc."""
    info = {"nproc": os.cpu_count()}
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name") and "model" not in info:
                info["model"] = line.split(":", 1)[1].strip()
            elif line.startswith("cpu MHz") and "mhz" not in info:
                info["mhz"] = float(line.split(":", 1)[1])
    except OSError:
        pass
    try:
        import subprocess

        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "CPU max MHz", "Socket(s)", "Core(s) per socket", "Thread(s) per core"):
                info["lscpu " + k.strip()] = v.strip()
    except Exception:
        pass
    return info


def cpu_threads_default():
    return max(1, min(os.cpu_count() or 1, CPU_SHARE))


def percentile(xs, p):
    return float(np.percentile(np.asarray(xs), p)) if len(xs) else None


def concurrent_streams(oa, torch, env, A, B, lvs, want, candidates=12, k=24):
    """Up to `want` HIP streams (default priority) whose batches overlap.

    HIP spreads a process's streams over a few hardware queues
    (GPU_MAX_HW_QUEUES, 4 unless set) by a least-used rule whose ties the
    program cannot see, and has no call that names a stream's queue; two
    streams on one queue run their kernels one after the other.  A pair is
    taken to overlap when k steps issued alternately on the two finish in
    under 0.8x the time of k steps on one (about 0.57x when they overlap,
    1.0x when they share a queue).  Candidates are kept greedily while each
    overlaps with every stream kept so far.  Returns (streams, pair checks)."""
    null = torch.cuda.current_stream()
    pool = [torch.cuda.Stream() for _ in range(candidates)]
    outs = [torch.empty(A.shape[1], dtype=torch.uint8, device=A.device) for _ in range(2)]
    for s in pool:  # each stream's survivor workspace is allocated on its first call
        oa.check_motions(env, A, B, lvs, out=outs[0], stream=s.cuda_stream)
    torch.cuda.synchronize()

    def timed(streams):
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(null)
        for s in streams:
            s.wait_event(ev0)
        for i in range(k):
            j = i % len(streams)
            oa.check_motions(env, A, B, lvs, out=outs[j], stream=streams[j].cuda_stream)
        for s in streams:
            null.wait_stream(s)
        ev1.record(null)
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1)

    one = min(timed([pool[0]]) for _ in range(2))
    kept, checks = [pool[0]], 0
    for c in pool[1:]:
        if len(kept) == want:
            break
        ok = True
        for s in kept:
            checks += 1
            if timed([s, c]) >= 0.8 * one:
                ok = False
                break
        if ok:
            kept.append(c)
    return kept, checks


def served_stream(oa, n_requests, qps, n_envs, plain_workers, seed=8, max_iters=2000, log_path=None,
                  check_parts=0, profile_slices=False, run_plain=True, block_threads=0, slice_us=0, prm_every=4,
                  prm_samples=2048):
    """Poisson stream of planning requests over seeded environments (16-1024
    obstacles, log-uniform, half spheres / half boxes): RRT-Connect, and every
    prm_every-th request a PRM query (batch roadmap of prm_samples samples),
    served (a) by the continuous-batching server and (b) by the plain path
    (host planner, one GPU launch + sync per check) on `plain_workers`
    threads.  Latency = completion - arrival.  Returns a dict of p50/p99
    latency and throughput, overall and per planner."""
    import concurrent.futures as cf
    import threading

    import torch

    from ompl_amd import serve

    rng = np.random.default_rng(seed)
    lvs = lvs_0p01()
    envs, starts = [], []
    for e in range(n_envs):
        n_obs = int(round(math.exp(rng.uniform(math.log(16), math.log(1024)))))
        sph, box = oa.workload.obstacles(1000 + e, n_obs - n_obs // 2, n_obs // 2)
        env = oa.Env(sph, box)
        q = oa.workload.configs(2000 + e, 4096)
        v = oa.validate_states(env, torch.from_numpy(q.T.copy()).cuda()).cpu().numpy()
        ok = q[v == 1]
        if len(ok) >= 2:
            envs.append(env)
            starts.append(ok)
    reqs = []
    for i in range(n_requests):
        k = i % len(envs)
        cand = starts[k]
        a, b = cand[rng.integers(len(cand))], cand[rng.integers(len(cand))]
        prm = prm_every > 0 and i % prm_every == prm_every - 1  # every prm_every-th request is a PRM query
        reqs.append(serve.make_request(envs[k], a, b, seed=10_000 + i, lvs=lvs, max_iters=max_iters,
                                       planner=serve.PRM if prm else serve.RRT_CONNECT, prm_samples=prm_samples))
    gaps = rng.exponential(1.0 / qps, n_requests)
    arrivals = np.cumsum(gaps)

    def pct(xs):
        return {"p50": float(np.percentile(xs, 50)), "p99": float(np.percentile(xs, 99)),
                "mean": float(np.mean(xs))}

    # (a) served (optionally the library's HIP events time each slice launch)
    if profile_slices:
        oa.profile.reset()
        oa.profile.enable(True)
    srv = serve.Server(hbm_budget_bytes=4 << 30, max_active=1024, max_nodes=4096, check_obstacle_parts=check_parts,
                       block_threads=block_threads, slice_us=slice_us)
    ids = [None] * n_requests
    t0 = time.perf_counter()
    for i, r in enumerate(reqs):
        dt = arrivals[i] - (time.perf_counter() - t0)
        if dt > 0:
            time.sleep(dt)
        ids[i] = srv.submit(r)
    res_s = [srv.wait(rid, 600_000) for rid in ids]
    span_s = time.perf_counter() - t0
    stats = srv.stats()
    lat_s = [r.latency_ms for r in res_s]
    paths_s = [srv.path(rid) for rid in ids]
    checks_s = sum(r.motion_checks for r in res_s)
    phases = srv.phase_us()
    srv.close()
    slice_us_gpu = None
    if profile_slices:
        n_sl, ms_sl = oa.profile.query("serve_slice")
        slice_us_gpu = 1e3 * ms_sl / n_sl if n_sl else None
        oa.profile.enable(False)
        oa.profile.reset()
    is_prm = [r.planner == serve.PRM for r in reqs]
    served = {"latency_ms": pct(lat_s), "throughput_rps": n_requests / span_s, "slices": stats["slices"],
              "latency_ms_rrt_connect": pct([x for x, q in zip(lat_s, is_prm) if not q]),
              "latency_ms_prm": pct([x for x, q in zip(lat_s, is_prm) if q]) if any(is_prm) else None,
              "grow_steps": stats["steps"], "motion_checks": checks_s,
              "edge_checks_per_s": checks_s / span_s,
              "run_ms": pct([r.run_ms for r in res_s]), "queued_ms": pct([r.queued_ms for r in res_s]),
              "device_us_per_step": {k: v / max(stats["steps"], 1) for k, v in phases.items()},
              "config": {"check_obstacle_parts": check_parts, "block_threads": block_threads or 512,
                         "slice_us": slice_us or 200}}
    if profile_slices:
        served["gpu_us_per_slice"] = slice_us_gpu
    if not run_plain:
        return {"requests": n_requests, "qps": qps, "served": served}

    # (b) plain: FIFO queue drained by plain_workers threads
    lat_p = [0.0] * n_requests
    res_p = [None] * n_requests
    lock = threading.Lock()

    def work(i, arrive):
        res, path = serve.plan_plain(reqs[i])
        with lock:
            lat_p[i] = (time.perf_counter() - arrive) * 1e3
            res_p[i] = (res, path)

    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=plain_workers) as ex:
        futs = []
        for i in range(n_requests):
            dt = arrivals[i] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            futs.append(ex.submit(work, i, t0 + arrivals[i]))
        for f in futs:
            f.result()
    span_p = time.perf_counter() - t0
    same = all(res_p[i][0].status == res_s[i].status and np.array_equal(res_p[i][1], paths_s[i])
               and res_p[i][0].motion_checks == res_s[i].motion_checks for i in range(n_requests))
    checks_p = sum(r.motion_checks for r, _ in res_p)
    solved = sum(r.status == serve.SOLVED for r in res_s)
    if log_path:
        from ompl_amd import benchlog

        common = {"range": reqs[0].range, "longest valid segment": lvs, "max iterations": max_iters}
        benchlog.write_log(log_path, "rrtconnect_served_vs_plain", [
            {"name": "geometric_RRTConnect_served", "common": common, "runs": benchlog.plan_runs(res_s, paths_s)},
            {"name": "geometric_RRTConnect_plain", "common": common,
             "runs": benchlog.plan_runs([r for r, _ in res_p], [p for _, p in res_p])}],
            seed=seed, total_duration=span_s + span_p,
            properties={"qps REAL": float(qps), "requests INTEGER": n_requests, "environments INTEGER": len(envs)})
    return {
        "requests": n_requests, "qps": qps, "envs_used": len(envs), "envs_drawn": n_envs,
        "solved": solved, "max_iters": max_iters,
        "served": served,
        "prm_requests": int(sum(is_prm)),
        "plain": {"latency_ms": pct(lat_p), "throughput_rps": n_requests / span_p, "workers": plain_workers,
                  "latency_ms_rrt_connect": pct([x for x, q in zip(lat_p, is_prm) if not q]),
                  "latency_ms_prm": pct([x for x, q in zip(lat_p, is_prm) if q]) if any(is_prm) else None,
                  "motion_checks": checks_p, "edge_checks_per_s": checks_p / span_p},
        "p99_plain_over_served": float(np.percentile(lat_p, 99) / np.percentile(lat_s, 99)),
        "p50_plain_over_served": float(np.percentile(lat_p, 50) / np.percentile(lat_s, 50)),
        "edge_checks_served_over_plain": float((checks_s / span_s) / (checks_p / span_p)),
        "identical_plans": bool(same),
    }


def prm_batch(oa, n_samples, cpu_samples, cpu_threads, run_cpu):
    """Batched PRM roadmap (samples -> validity -> kNN k=10 -> candidate edges ->
    motion checks -> components) on 200k samples: the 512-sphere + 512-box
    scene (empty roadmap: no sample is collision free), the same count of
    small obstacles, and the 64-sphere scene of the validity batch, where most
    samples are valid.  The CPU restatement runs on a bounded sample."""
    from ompl_amd.prm import build_roadmap

    out = {}
    scenes = (("1024_obstacles", 1003, 512, 512, (0.05, 0.15), (0.05, 0.2)),
              # the same count of small obstacles (about 9 % of samples collision free): exercises kNN, motion
              # checks and graph assembly at 1,024 obstacles, which the scene above leaves empty
              ("1024_small_obstacles", 1003, 512, 512, (0.01, 0.03), (0.01, 0.03)),
              ("64_spheres", 1000, 64, 0, (0.05, 0.15), (0.05, 0.2)))
    for name, env_seed, ns, nb, rr, hh in scenes:
        s, b = oa.workload.obstacles(env_seed, ns, nb, r=rr, h=hh)
        env = oa.Env(s, b)
        # warm the pipeline and tune the motion launch shape on this roadmap's candidate edges
        build_roadmap(env, PRM_SEED, n_samples, n_streams=256, fetch=False, autotune=True)
        tuning = {k: v for k, v in env.tuning().items() if k != "own"}  # the environment's own shapes
        reps = []
        for _ in range(3):
            t0 = time.perf_counter()
            st = build_roadmap(env, PRM_SEED, n_samples, n_streams=256, fetch=False)
            reps.append(time.perf_counter() - t0)
        best = min(reps)
        row = {"samples": st["samples"], "milestones": st["milestones"], "candidate_edges": st["candidate_edges"],
               "valid_edges": st["valid_edges"], "components": st["components"], "build_ms": 1e3 * best,
               "stage_ms": {k[3:]: round(st[k], 3) for k in st if k.startswith("ms_")},
               "edge_checks_per_s": st["candidate_edges"] / (st["ms_motions"] * 1e-3) if st["ms_motions"] > 0 else None,
               "motion_tuning": tuning}
        if run_cpu:
            from oracle import oracle_py

            oenv = oracle_py.OracleEnv(s, b)
            t0 = time.perf_counter()
            r = oracle_py.prm_batch(oenv, PRM_SEED, cpu_samples, 10, lvs_0p01(), 1, 256, threads=cpu_threads)
            cpu_s = time.perf_counter() - t0
            g = build_roadmap(env, PRM_SEED, cpu_samples, n_streams=256)
            same = (np.array_equal(g.sample_index, r["sample_index"]) and np.array_equal(g.edges, r["edges"])
                    and np.array_equal(g.edge_valid, r["edge_valid"]) and np.array_equal(g.component, r["component"]))
            row["cpu_baseline"] = {"samples_per_s": r["samples"].shape[0] / cpu_s, "cores": cpu_threads,
                                   "kind": "port", "sample": f"{r['samples'].shape[0]} samples, {cpu_s:.1f} s"}
            row["gpu_samples_per_s"] = st["samples"] / best
            row["parity_bit_exact_on_sample"] = bool(same)
        out[name] = row
    return out


TRAFFIC_RECORD = ROOT / "profiles" / "motion_batch_16384_edges_256_obstacles.hbm_traffic.json"


def hbm_traffic(kernel: str, tuning: dict):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE records, when one was taken with this launch
    shape; otherwise None."""
    try:
        rec = json.loads(TRAFFIC_RECORD.read_text())
    except (OSError, ValueError):
        return None, None
    keys = ("validity_lanes_per_state", "validity_obstacle_parts", "motion_lanes_per_state",
            "motion_obstacle_parts", "motion_waves_per_cu", "motion_strategy", "validity_shared_fk")
    # the kernels that run for this tuning (shared FK exists for (8, parts > 1) only)
    lw, ow, le, oe = (tuning["motion_lanes_per_state"], tuning["motion_obstacle_parts"],
                      tuning["validity_lanes_per_state"], tuning["validity_obstacle_parts"])
    shared = "true" if tuning["motion_strategy"] == 4 and lw == 8 and ow > 1 else "false"
    if kernel == "motion_chunk":  # both phases (the timed batches ask for no checked counts: no results kernel)
        names = [f"motion_chunk_kernel<{lw}, {ow}, false, {shared}, {ph}>" for ph in (1, 2)]
    elif kernel == "motion_endpoint":
        es = tuning.get("validity_shared_fk") and le == 8 and oe > 1
        names = [f"{'motion_endpoint_shared_kernel' if es else 'motion_endpoint_kernel'}<{le}, {oe}>"]
    else:
        return None, None
    # a record taken with this tuning first; else any record of the same kernels
    # (their sweep dispatches), noted as such
    for same in (True, False):
        for r in rec["records"]:
            if same and any(r["tuning"].get(k) != tuning.get(k) for k in keys):
                continue
            found = {n: (k, v) for n in names for k, v in r["bytes_per_dispatch"].items() if n + "(" in k}
            if len(found) == len(names):
                src = " + ".join(k for k, _ in found.values())
                note = "" if same else ", recorded under another tuning"
                return int(sum(v for _, v in found.values())), f"{TRAFFIC_RECORD.name} ({src}{note})"
    return None, f"{TRAFFIC_RECORD.name}: no record for this launch shape"


def timed(fn, reps, warm=2):
    """Average milliseconds of fn() over reps launches (HIP events on torch's
    current stream, which is the stream the library launches on)."""
    import torch

    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def gnat_baseline(oracle_py, metric, db, qs, k, radius, threads, gpu_idx, seconds=1.5):
    """NearestNeighborsGNAT (the planners' default structure; oracle restatement,
    FP64 distances) on the host: bulk build over the whole database, then
    k-NN (and radius) queries in chunks until `seconds` have passed.  The
    GPU's FP32 index lists agree with GNAT's FP64 ones except at near-ties;
    the agreeing fraction of the timed queries is reported."""
    t0 = time.perf_counter()
    g = oracle_py.Gnat(metric, db, 1)
    build_s = time.perf_counter() - t0
    # chunks of the workload's queries, cycling over them, until `seconds`
    # have passed (the agreement is taken over the first pass)
    done, agree, checked, t0 = 0, 0, 0, time.perf_counter()
    chunk = max(threads * 4, 16)
    while time.perf_counter() - t0 < seconds:
        at = done % len(qs)
        gi, _ = g.knn(qs[at:at + chunk], k, threads)
        if done < len(qs):
            agree += int(np.sum(np.all(gi == gpu_idx[at:at + chunk], axis=1)))
            checked += len(gi)
        done += len(gi)
    knn_s = time.perf_counter() - t0
    out = {"queries_per_s": done / knn_s, "cores": threads, "kind": "port", "build_s": build_s,
           "sample": f"GNAT (degree 8, leaf 50) over all {len(db):,} points: {done} k-NN queries (cycling over the "
                     f"{len(qs):,}) in {knn_s:.1f} s after a {build_s:.1f} s build; {CPU_FLAGS}",
           "knn_index_lists_equal_fraction": agree / max(checked, 1)}
    if radius is not None:
        done_r, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < seconds:
            at = done_r % len(qs)
            g.radius(qs[at:at + chunk], radius, threads)
            done_r += len(qs[at:at + chunk])
        rs = time.perf_counter() - t0
        out["radius_queries_per_s"] = done_r / rs
        out["radius_sample"] = f"{done_r} radius queries (cycling over the {len(qs):,}) in {rs:.1f} s"
    g.close()
    return out


def scope_rows(oa, run_cpu, cpu_threads, only=None, sparse_tuning=None):
    """Throughput of the other rows of the hot path, each on its configuration,
    with its algorithmic work and a bounded CPU-restatement baseline.  only:
    names of the rows to run (None: all)."""
    import torch

    from ompl_amd.nn import NearestNeighborsGPU
    from ompl_amd.sampler import sample_uniform, sample_valid

    rows = {}
    oracle_py = None
    if run_cpu:
        from oracle import oracle_py

    def roof(flops, ms, peak=PEAK_FP32_TFLOPS):
        a = flops / (ms * 1e-3) / 1e12
        return {"achieved_tflops": a, "frac": a / peak, "algorithmic_gflop": flops / 1e9}

    def want(name):
        return only is None or name in only

    # validity: 65,536 configurations U[-pi,pi]^7 against 64 spheres
    if want("validity_64_spheres"):
        validity_row(oa, rows, oracle_py, cpu_threads, roof, timed)
    if want("knn_l2_100k"):
        knn_row(oa, rows, oracle_py, cpu_threads, roof, timed, NearestNeighborsGPU)
    if want("se3_1m"):
        se3_row(oa, rows, oracle_py, cpu_threads, roof, timed, NearestNeighborsGPU)
    if want("sampling"):
        sampling_rows(oa, rows, oracle_py, timed, sample_uniform, sample_valid)
    if want("state_lists"):
        state_list_row(oa, rows, oracle_py, timed, roof)
    if want("motion_sparse"):
        motion_sparse_row(oa, rows, oracle_py, cpu_threads, roof, timed, sparse_tuning)
    if want("last_valid"):
        last_valid_row(oa, rows, oracle_py, timed, roof)
    return rows


def sequential_motion_states(a_rows, b_rows, lvs, valid, last_t):
    """The states DiscreteMotionValidator::checkMotion(s1, s2, lastValid)
    checks, in its order (positions 1..nd-1, then s2 as position nd), up to the
    first invalid one (every position of a valid edge): FP32 [7, n], for the
    roofline's reference-order work.  nd and the interpolation are the
    library's (FP64, then FP32)."""
    diff = a_rows - b_rows
    d = np.sqrt(np.sum(diff * diff, axis=1))
    nd = np.ceil(d / lvs).astype(np.int64)
    last = np.maximum(nd, 1)
    # lastValid (j-1)/nd for the first invalid position j (s2: (nd-1)/nd)
    j = np.where(valid == 1, last, np.rint(np.nan_to_num(last_t) * nd).astype(np.int64) + 1)
    j = np.minimum(j, last)
    edge = np.repeat(np.arange(len(a_rows)), j)
    pos = np.arange(int(j.sum())) - np.repeat(np.cumsum(j) - j, j) + 1
    t = pos / np.maximum(nd[edge], 1)
    st = a_rows[edge] + (b_rows[edge] - a_rows[edge]) * t[:, None]
    s2 = pos >= last[edge]
    st[s2] = b_rows[edge[s2]]
    return st.astype(np.float32).T.copy()


def last_valid_row(oa, rows, oracle_py, timed, roof):
    """DiscreteMotionValidator::checkMotion(s1, s2, lastValid) batched over the
    16,384 edges of the dense batch (128 + 128 obstacles) and of the sparse
    batch (16 + 16 obstacles, lognormal(-0.5, 1) lengths), with the launch
    shape in force."""
    import torch

    out = {}
    for name, seed, ns, nb, eseed, mu in (("dense_256_obstacles", ENV_SEED, N_SPHERES, N_BOXES, EDGE_SEED, -1.0),
                                          ("sparse_32_obstacles", 1002, 16, 16, 9, -0.5)):
        s, b = oa.workload.obstacles(seed, ns, nb)
        env = oa.Env(s, b)
        a_rows, b_rows = oa.workload.edges(eseed, N_EDGES, mu, 1.0, 0.05, 5.0)
        lvs = lvs_0p01()
        A = torch.from_numpy(a_rows.T.copy()).cuda()
        B = torch.from_numpy(b_rows.T.copy()).cuda()
        ms = timed(lambda: oa.check_motions_last_valid(env, A, B, lvs), 20)
        v, t = oa.check_motions_last_valid(env, A, B, lvs)
        v, t = v.cpu().numpy(), t.cpu().numpy()
        w = oa.validity_work(env, torch.from_numpy(sequential_motion_states(a_rows, b_rows, lvs, v, t)).cuda())
        r = {"edges": N_EDGES, "obstacles": ns + nb, "ms": ms, "edge_checks_per_s": N_EDGES / (ms * 1e-3),
             "valid_fraction": float(v.mean()), "states_checked": w["states"],
             "roofline": roof(oa.work_flops(w), ms)}
        if ns + nb < 64:
            # the scene's own launch shapes (the motion autotuner on the same edges)
            tuned, _ = env.autotune(A, B, lvs, reps=3)
            ms_own = timed(lambda: oa.check_motions_last_valid(env, A, B, lvs), 20)
            v2, t2 = oa.check_motions_last_valid(env, A, B, lvs)
            r["own_shape"] = {"ms": ms_own, "edge_checks_per_s": N_EDGES / (ms_own * 1e-3), "tuning": tuned,
                              "roofline": roof(oa.work_flops(w), ms_own),
                              "same_result": bool(np.array_equal(v2.cpu().numpy(), v)) and
                              bool(np.array_equal(t2.cpu().numpy(), t, equal_nan=True))}
        if oracle_py is not None:
            oenv = oracle_py.OracleEnv(s, b)
            t0 = time.perf_counter()
            same = True
            for i in range(2048):
                ov, ot, _ = oracle_py.check_motion_last_valid(oenv, a_rows[i], b_rows[i], lvs)
                same &= bool(v[i]) == bool(ov) and (bool(ov) or t[i] == ot)
            dt = time.perf_counter() - t0
            r["cpu_baseline"] = {"edge_checks_per_s": 2048 / dt, "cores": 1, "kind": "port",
                                 "sample": f"2,048 of the 16,384 edges, {dt:.2f} s"}
            r["parity_bit_exact_on_sample"] = bool(same)
        env.close()
        out[name] = r
    rows["last_valid"] = out


def motion_sparse_row(oa, rows, oracle_py, cpu_threads, roof, timed, own_tuning=None):
    """The motion check where most edges are collision-free: 16 spheres + 16
    boxes, 16,384 edges of lognormal(-0.5, 1) length, so that most edges pass
    their end state and are walked to full resolution (the walk, not the end
    check, carries the time).  Timed with the launch shape in force (the
    dense batch's) and with the environment's own shape, which the
    autotuner picks for this batch (or `own_tuning` when given)."""
    import torch

    s, b = oa.workload.obstacles(1002, 16, 16)
    env = oa.Env(s, b)
    a_rows, b_rows = oa.workload.edges(9, N_EDGES, -0.5, 1.0, 0.05, 5.0)
    lvs = lvs_0p01()
    A = torch.from_numpy(a_rows.T.copy()).cuda()
    B = torch.from_numpy(b_rows.T.copy()).cuda()
    out = torch.empty(N_EDGES, dtype=torch.uint8, device="cuda")
    work = oa.motion_work(env, A, B, lvs)
    flops = oa.work_flops(work)
    flops_end = oa.work_flops(oa.validity_work(env, B.float().contiguous()))

    def measure():
        ms = timed(lambda: oa.check_motions(env, A, B, lvs, out=out), 50, warm=5)
        valid = out.cpu().numpy().copy()
        oa.profile.reset()
        oa.profile.enable(True)
        for _ in range(20):
            oa.check_motions(env, A, B, lvs, out=out)
        torch.cuda.synchronize()
        kernels = {}
        for name in ("motion_fused", "motion_endpoint", "motion_walk", "motion_chunk"):
            n, kms = oa.profile.query(name)
            if n:
                kernels[name] = {"launches": n, "avg_us": 1e3 * kms / n}
        oa.profile.enable(False)
        oa.profile.reset()
        m = {"ms": ms, "edge_checks_per_s": N_EDGES / (ms * 1e-3), "states_checked_per_s": work["states"] / (ms * 1e-3),
             "kernels": kernels, "roofline_step": roof(flops, ms)}
        walk = next((k for k in ("motion_walk", "motion_chunk") if k in kernels), None)
        if walk is not None:
            r = roof(flops - flops_end, kernels[walk]["avg_us"] * 1e-3)
            m["roofline_walk"] = dict(r, kernel=walk, kernel_avg_us=kernels[walk]["avg_us"])
        return m, valid

    saved = oa.get_tuning()
    dense, valid = measure()
    if own_tuning:
        tuned = env.set_tuning(own_tuning)
    else:
        tuned, _ = env.autotune(A, B, lvs, reps=5)  # the environment's own shapes; the library's stay
    own, valid_own = measure()
    row = {"edges": N_EDGES, "obstacles": 32, "states_per_step": work["states"], "valid_fraction": float(valid.mean()),
           "algorithmic_gflop_per_step": flops / 1e9, "dense_batch_shape": dict(dense, tuning=saved),
           "own_shape": dict(own, tuning=tuned)}
    if oracle_py is not None:
        oenv = oracle_py.OracleEnv(s, b)
        t0 = time.perf_counter()
        ref, _ = oracle_py.check_motions(oenv, a_rows, b_rows, lvs, threads=cpu_threads)
        dt = time.perf_counter() - t0
        row["cpu_baseline"] = {"edge_checks_per_s": N_EDGES / dt, "cores": cpu_threads, "kind": "port",
                               "sample": f"the full 16,384-edge batch, {dt:.2f} s; {CPU_FLAGS}"}
        row["parity_bit_exact"] = bool(np.array_equal(ref, valid)) and bool(np.array_equal(ref, valid_own))
    env.close()
    rows["motion_sparse"] = row


def validity_row(oa, rows, oracle_py, cpu_threads, roof, timed):
    import torch

    s, b = oa.workload.obstacles(1000, 64, 0)
    env = oa.Env(s, b)
    q = oa.workload.configs(1, 65536)
    qd = torch.from_numpy(q.T.astype(np.float32).copy()).cuda()
    out = torch.empty(65536, dtype=torch.uint8, device="cuda")
    shapes = {}
    saved = oa.get_tuning()
    for lps, ops, shared in ((8, 1, 0), (8, 2, 0), (8, 8, 0), (8, 2, 1), (8, 4, 1), (8, 8, 1), (7, 1, 0), (7, 2, 0), (7, 4, 0), (7, 9, 0), (1, 1, 0)):
        oa.set_tuning(validity_lanes_per_state=lps, validity_obstacle_parts=ops, validity_shared_fk=shared)
        shapes[f"{lps}x{ops}{'-shared' if shared else ''}"] = timed(lambda: oa.validate_states(env, qd, out=out), 50)
    best = min(shapes, key=shapes.get)
    ms = shapes[best]
    oa.set_tuning(**saved)
    oa.validate_states(env, qd, out=out)
    w = oa.validity_work(env, qd)
    row = {"states": 65536, "obstacles": 64, "ms": ms, "states_per_s": 65536 / (ms * 1e-3), "shape": best,
           "ms_by_shape": shapes, "valid_fraction": float(out.float().mean().item()),
           "roofline": roof(oa.work_flops(w), ms)}
    if oracle_py is not None:
        oenv = oracle_py.OracleEnv(s, b)
        t0 = time.perf_counter()
        ref = oracle_py.validate_states_mt(oenv, q.T.astype(np.float32), cpu_threads)
        row["cpu_baseline"] = {"states_per_s": 65536 / (time.perf_counter() - t0), "cores": cpu_threads,
                               "kind": "port", "sample": "the full 65,536-state batch"}
        row["parity_bit_exact"] = bool(np.array_equal(ref, out.cpu().numpy()))
    rows["validity_64_spheres"] = row


def knn_row(oa, rows, oracle_py, cpu_threads, roof, timed, NearestNeighborsGPU):
    import torch

    # k-NN, R^7 L2: 100,000 points, 4,096 queries, k = 10
    db = oa.workload.configs(2, 100_000).astype(np.float32)
    qs = oa.workload.configs(3, 4096).astype(np.float32)
    nn = NearestNeighborsGPU(oa.METRIC_L2, 7)
    nn.add(torch.from_numpy(db).cuda())
    qsd = torch.from_numpy(qs).cuda()
    oi = torch.empty((4096, 10), dtype=torch.int32, device="cuda")
    od = torch.empty((4096, 10), dtype=torch.float32, device="cuda")
    ms = timed(lambda: nn.knn_dev(qsd, 10, oi, od), 20)
    row = {"db": 100_000, "queries": 4096, "k": 10, "ms": ms, "queries_per_s": 4096 / (ms * 1e-3),
           "roofline": roof(20.0 * 100_000 * 4096, ms)}
    # packed batch: 16 concurrent requests of 64..512 queries with k in
    # {1, 4, 10, 16} in one launch, against one launch per request
    prng = np.random.default_rng(17)
    sizes = prng.integers(64, 513, 16)
    kreq = prng.choice([1, 4, 10, 16], 16)
    offs = np.concatenate([[0], np.cumsum(sizes)])
    pq = torch.from_numpy(qs[np.arange(offs[-1]) % 4096]).cuda()
    pk = torch.from_numpy(np.repeat(kreq, sizes).astype(np.int32)).cuda()
    ms_packed = timed(lambda: nn.knn_packed_dev(pq, pk, 16, trim=False), 20)
    parts = [(pq[offs[i]:offs[i + 1]], int(kreq[i])) for i in range(16)]
    ms_each = timed(lambda: [nn.knn_dev(qp, kp) for qp, kp in parts], 20)
    row["packed_requests"] = {"requests": 16, "queries": int(offs[-1]), "k_per_request": [int(v) for v in kreq],
                              "ms_one_launch": ms_packed, "ms_one_launch_per_request": ms_each,
                              "speedup": ms_each / ms_packed}
    if oracle_py is not None:
        t0 = time.perf_counter()
        ri, _ = oracle_py.knn_mt(0, db, qs[:512], 10, cpu_threads)
        dt = time.perf_counter() - t0
        row["cpu_linear"] = {"queries_per_s": 512 / dt, "cores": cpu_threads, "kind": "port",
                             "sample": f"512 of the 4,096 queries, {dt:.1f} s, FP32 brute force"}
        row["parity_bit_exact_on_sample"] = bool(np.array_equal(ri, oi[:512].cpu().numpy()))
        row["cpu_baseline"] = gnat_baseline(oracle_py, 0, db, qs, 10, None, cpu_threads, oi.cpu().numpy())
    nn.close()
    rows["knn_l2_100k"] = row


def se3_row(oa, rows, oracle_py, cpu_threads, roof, timed, NearestNeighborsGPU):
    import torch

    # SE3 at scale: 1,000,000 points, 65,536 queries, k = 16 and the RRT* radius
    db = oa.workload.se3(5, 1_000_000)
    qs = oa.workload.se3(6, 65536)
    nn = NearestNeighborsGPU(oa.METRIC_SE3, 7)
    nn.add(torch.from_numpy(db).cuda())
    qsd = torch.from_numpy(qs).cuda()
    oi = torch.empty((65536, 16), dtype=torch.int32, device="cuda")
    od = torch.empty((65536, 16), dtype=torch.float32, device="cuda")
    ms_k = timed(lambda: nn.knn_dev(qsd, 16, oi, od), 3, warm=1)
    r = oa.rrtstar_radius(6, 8000.0 * np.pi ** 2, 1_000_000)
    ms_r = timed(lambda: nn.radius_dev(qsd, r), 3, warm=1)
    off, _, _ = nn.radius_dev(qsd, r)
    pairs = 65536.0 * 1_000_000
    row = {"db": 1_000_000, "queries": 65536, "k": 16, "radius": r, "knn_ms": ms_k,
           "knn_queries_per_s": 65536 / (ms_k * 1e-3), "radius_ms": ms_r,
           "radius_queries_per_s": 65536 / (ms_r * 1e-3), "radius_mean_hits": float(off[-1].item()) / 65536,
           "roofline_knn": roof(20.0 * pairs, ms_k), "roofline_radius": roof(20.0 * pairs, ms_r),
           "work_note": "brute-force pairs x 20 FLOP-eq (acos and sqrt counted as 1)"}
    if oracle_py is not None:
        t0 = time.perf_counter()
        ri, _ = oracle_py.knn_mt(1, db, qs[:64], 16, cpu_threads)
        dt = time.perf_counter() - t0
        row["cpu_linear"] = {"queries_per_s": 64 / dt, "cores": cpu_threads, "kind": "port",
                             "sample": f"64 of the 65,536 k-NN queries, {dt:.1f} s, FP32 brute force"}
        row["parity_bit_exact_on_sample"] = bool(np.array_equal(ri, oi[:64].cpu().numpy()))
        row["cpu_baseline"] = gnat_baseline(oracle_py, 1, db, qs, 16, r, cpu_threads, oi.cpu().numpy())
    nn.close()
    rows["se3_1m"] = row


def sampling_rows(oa, rows, oracle_py, timed, sample_uniform, sample_valid):
    # uniform sampling: 256 generators x 782 states (200,192 states)
    lo, hi = np.full(7, -np.pi), np.full(7, np.pi)
    seeds = np.arange(256, dtype=np.uint32) + PRM_SEED
    ms = timed(lambda: sample_uniform(seeds, 782, lo, hi, f32_soa=True), 10)
    n = 256 * 782
    rows["sample_uniform"] = {"states": n, "ms": ms, "states_per_s": n / (ms * 1e-3),
                              "bytes_written_per_state": 84,
                              "hbm_gbps": n * 84 / (ms * 1e-3) / 1e9,
                              "note": "host-staged seeds and bounds and a stream sync are inside the time"}
    if oracle_py is not None:
        t0 = time.perf_counter()
        for sd in seeds[:32].tolist():
            oracle_py.Mt19937(sd).sample_uniform(lo, hi, 782)
        dt = time.perf_counter() - t0
        rows["sample_uniform"]["cpu_baseline"] = {"states_per_s": 32 * 782 / dt, "cores": 1, "kind": "port",
                                                  "sample": f"32 of the 256 generators, {dt:.2f} s"}

    # valid sampling: 256 samplers x 64 calls in the 128 + 128 obstacle scene (1.2% valid)
    s2, b2 = oa.workload.obstacles(ENV_SEED, N_SPHERES, N_BOXES)
    env2 = oa.Env(s2, b2)
    ms = timed(lambda: sample_valid(env2, seeds, 64), 3, warm=1)
    _, found, draws = sample_valid(env2, seeds, 64)
    rows["sample_valid"] = {"samplers": 256, "calls": 64, "attempts": 100, "ms": ms,
                            "calls_per_s": 256 * 64 / (ms * 1e-3), "found_fraction": float(found.float().mean()),
                            "draws_per_s": float(draws.sum().item()) / (ms * 1e-3)}
    if oracle_py is not None:
        oenv2 = oracle_py.OracleEnv(s2, b2)
        t0 = time.perf_counter()
        for sd in seeds[:8].tolist():
            oracle_py.Mt19937(sd).sample_valid(oenv2, 64, 100)
        dt = time.perf_counter() - t0
        rows["sample_valid"]["cpu_baseline"] = {"calls_per_s": 8 * 64 / dt, "cores": 1, "kind": "port",
                                                "sample": f"8 of the 256 samplers, {dt:.2f} s"}


def state_list_row(oa, rows, oracle_py, timed, roof):
    import torch

    # precomputed state lists: the motion states of the 16,384 benchmark edges, first 64 each
    s2, b2 = oa.workload.obstacles(ENV_SEED, N_SPHERES, N_BOXES)
    env2 = oa.Env(s2, b2)
    a_rows, b_rows = oa.workload.edges(EDGE_SEED, 4096)
    t = np.linspace(0.0, 1.0, 64)
    st = (a_rows[:, None, :] + (b_rows - a_rows)[:, None, :] * t[None, :, None]).reshape(-1, 7)
    std = torch.from_numpy(st.T.copy()).cuda()
    offs = torch.arange(0, 4097 * 64, 64, dtype=torch.int64, device="cuda")
    ms = timed(lambda: oa.check_state_lists(env2, std, offs), 20)
    v, f = oa.check_state_lists(env2, std, offs)
    # reference-order work of the sequential overload: each list's states up
    # to its first invalid one
    fh = f.cpu().numpy()
    cnt = np.where(fh >= 0, fh + 1, 64)
    idx = np.repeat(np.arange(4096) * 64, cnt) + (np.arange(int(cnt.sum())) - np.repeat(np.cumsum(cnt) - cnt, cnt))
    w = oa.validity_work(env2, torch.from_numpy(st[idx].astype(np.float32).T.copy()).cuda())
    rows["state_lists"] = {"lists": 4096, "states_per_list": 64, "ms": ms, "lists_per_s": 4096 / (ms * 1e-3),
                           "valid_fraction": float(v.float().mean().item()), "states_checked": w["states"],
                           "roofline": roof(oa.work_flops(w), ms)}
    if oracle_py is not None:
        oenv2 = oracle_py.OracleEnv(s2, b2)
        t0 = time.perf_counter()
        same = True
        vh, fh = v.cpu().numpy(), f.cpu().numpy()
        for i in range(512):
            ok, first = oracle_py.check_states_first_invalid(oenv2, st[64 * i:64 * (i + 1)])
            same &= bool(ok) == bool(vh[i]) and (first if not ok else -1) == int(fh[i])
        dt = time.perf_counter() - t0
        rows["state_lists"]["cpu_baseline"] = {"lists_per_s": 512 / dt, "cores": 1, "kind": "port",
                                               "sample": f"512 of the 4,096 lists, {dt:.2f} s"}
        rows["state_lists"]["parity_bit_exact_on_sample"] = bool(same)
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--streams", type=int, default=4,
                    help="HIP streams the timed steps are issued on, round-robin (independent batches); streams "
                         "whose batches do not overlap (one hardware queue) are left out")
    ap.add_argument("--plain-edges", type=int, default=2048)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tuning", type=str, default=None, help="JSON dict of launch knobs (skips the autotuner)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the CPU baselines (0 = os.cpu_count, at most the 16-core share of one GPU)")
    ap.add_argument("--stream-requests", type=int, default=256, help="served-stream requests (0 = skip)")
    ap.add_argument("--stream-qps", type=float, default=200.0)
    ap.add_argument("--stream-envs", type=int, default=32)
    ap.add_argument("--plain-workers", type=int, default=8,
                    help="threads of the plain per-call planner on the served stream (0 = skip the plain leg)")
    ap.add_argument("--no-stream-sweeps", action="store_true",
                    help="skip the served stream's offered-rate and server-setting sweeps")
    ap.add_argument("--prm-samples", type=int, default=PRM_SAMPLES, help="PRM roadmap batch size (0 = skip)")
    ap.add_argument("--prm-cpu-samples", type=int, default=20000)
    ap.add_argument("--no-rows", action="store_true", help="skip the per-row measurements")
    ap.add_argument("--rows", type=str, default=None,
                    help="comma-separated rows to measure (validity_64_spheres, knn_l2_100k, se3_1m, sampling, "
                         "state_lists, motion_sparse, last_valid; default all)")
    ap.add_argument("--sparse-tuning", type=str, default=None,
                    help="JSON dict: the sparse-scene row's own launch shapes (skips its autotuner run)")
    ap.add_argument("--rows-only", action="store_true",
                    help="measure only the per-row workloads (with --tuning's shapes, else the library's) and print "
                         "them as one JSON line; no headline value")
    ap.add_argument("--benchmark-log", type=str, default=None,
                    help="also write the served stream's runs as an OMPL benchmark log file")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="only run 5 device copies of this many bytes (HBM counter calibration) and exit")
    args = ap.parse_args(argv)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); "
              "the multi-GPU sharding plan is described, not run", file=sys.stderr)
        return 2

    import torch

    if args.calibrate_copy:
        n = args.calibrate_copy // 4
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 5}))
        return 0

    import ompl_amd as oa

    oa.init(0)
    if args.rows_only:
        if args.tuning:
            oa.set_tuning(**json.loads(args.tuning))
        rows_res = scope_rows(oa, not args.no_cpu, args.cpu_threads or cpu_threads_default(),
                              only=set(args.rows.split(",")) if args.rows else None,
                              sparse_tuning=json.loads(args.sparse_tuning) if args.sparse_tuning else None)
        print(json.dumps({"rows": rows_res, "tuning": oa.get_tuning()}), flush=True)
        return 0
    spheres, boxes = oa.workload.obstacles(ENV_SEED, N_SPHERES, N_BOXES)
    env = oa.Env(spheres, boxes)
    a_rows, b_rows = oa.workload.edges(EDGE_SEED, N_EDGES)
    lvs = lvs_0p01()
    dev = torch.device("cuda", 0)
    A = torch.from_numpy(a_rows.T.copy()).to(dev)
    B = torch.from_numpy(b_rows.T.copy()).to(dev)
    out = torch.empty(N_EDGES, dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()

    # ---- launch shape: the project autotuner (or a fixed shape from --tuning)
    if args.tuning:
        oa.set_tuning(**json.loads(args.tuning))
        tuned_us = None
    else:
        _, tuned_us = oa.autotune_motions(env, A, B, lvs, reps=20)
    t = oa.get_tuning()

    # ---- the streams the timed steps are issued on: ones whose batches overlap
    picked, pair_checks = concurrent_streams(oa, torch, env, A, B, lvs, max(1, args.streams))
    # ---- the launch shape for batches on those streams (the one-batch-at-a-time
    #      shape above keeps more persistent walk waves than is best when the
    #      streams' batches share the device)
    t_one = t
    streams_us = None
    if not args.tuning and len(picked) > 1:
        t, streams_us = oa.autotune_motions_streams(env, A, B, lvs, picked)

    # ---- W untimed warmup steps, then the timed region: K batched steps,
    #      issued round-robin on S HIP streams (independent batches: each
    #      stream has its own output and survivor workspace), between events
    #      on the first stream, which waits for the others at the end
    def run_steps(n_streams, k, warm):
        use = picked[:n_streams]
        handles = [s.cuda_stream for s in use]
        outs = [out] + [torch.empty_like(out) for _ in use[1:]]
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        for i in range(warm):
            oa.check_motions(env, A, B, lvs, out=outs[i % len(use)], stream=handles[i % len(use)])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev0.record(use[0])
        for s in use[1:]:
            s.wait_event(ev0)
        for i in range(k):
            j = i % len(use)
            oa.check_motions(env, A, B, lvs, out=outs[j], stream=handles[j])
        for s in use[1:]:
            use[0].wait_stream(s)
        ev1.record(use[0])
        torch.cuda.synchronize()
        wall_s = time.perf_counter() - t0
        same = all(np.array_equal(o.cpu().numpy(), outs[0].cpu().numpy()) for o in outs[1:])
        return ev0.elapsed_time(ev1), wall_s, same

    step_ms = []
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    gpu_ms, wall, streams_same = run_steps(len(picked), args.steps, args.warmup)
    ms_per_step = gpu_ms / args.steps
    value = N_EDGES * args.steps / (gpu_ms / 1e3)
    valid = out.cpu().numpy().copy()
    pair_ms, _, _ = run_steps(2, args.steps, args.warmup) if len(picked) > 2 else (gpu_ms, wall, True)
    # one stream with the shape tuned for one stream
    oa.set_tuning(**t_one)
    single_ms, _, _ = run_steps(1, args.steps, args.warmup) if len(picked) > 1 else (gpu_ms, wall, True)
    oa.set_tuning(**t)

    # ---- per-step latency (synchronous steps) for p50/p99
    for _ in range(min(args.steps, 200)):
        s0 = time.perf_counter()
        oa.check_motions(env, A, B, lvs, out=out)
        torch.cuda.synchronize()
        step_ms.append((time.perf_counter() - s0) * 1e3)

    # ---- a few fixed launch shapes beside the tuned one (per-kernel times)
    sweep = {}
    for name, kw in (("s3_e8x8_w8x8", dict(motion_strategy=3, validity_obstacle_parts=8, motion_obstacle_parts=8)),
                     ("s3_e8x8sh_w8x4", dict(motion_strategy=3, validity_obstacle_parts=8, motion_obstacle_parts=4,
                                             validity_shared_fk=1)),
                     ("s4_e8x8_w8x8", dict(motion_strategy=4, validity_obstacle_parts=8, motion_obstacle_parts=8)),
                     ("s4_e8x8sh_w8x2", dict(motion_strategy=4, validity_obstacle_parts=8, motion_obstacle_parts=2,
                                             validity_shared_fk=1)),
                     ("s3_e7x9_w7x9", dict(motion_strategy=3, validity_lanes_per_state=7, validity_obstacle_parts=9,
                                           motion_lanes_per_state=7, motion_obstacle_parts=9)),
                     ("s3_e8x8_w7x9", dict(motion_strategy=3, validity_obstacle_parts=8, motion_lanes_per_state=7,
                                           motion_obstacle_parts=9)),
                     ("s3_e7x4_w7x9", dict(motion_strategy=3, validity_lanes_per_state=7, validity_obstacle_parts=4,
                                           motion_lanes_per_state=7, motion_obstacle_parts=9)),
                     ("s3_e7x2_w7x9", dict(motion_strategy=3, validity_lanes_per_state=7, validity_obstacle_parts=2,
                                           motion_lanes_per_state=7, motion_obstacle_parts=9))):
        oa.set_tuning(**{**t, "validity_lanes_per_state": 8, "motion_lanes_per_state": 8, "validity_shared_fk": 0,
                         **kw})
        oa.profile.reset()
        oa.profile.enable(True)
        ms = timed(lambda: oa.check_motions(env, A, B, lvs, out=out), 30)
        oa.profile.enable(False)
        per = {}
        for kname in ("motion_endpoint", "motion_chunk"):
            n, kms = oa.profile.query(kname)
            if n:
                per[kname] = 1e3 * kms / n
        oa.profile.reset()
        sweep[name] = {"ms_per_step": ms, "kernel_us": per}
    oa.set_tuning(**t)

    # ---- the same batch replayed from a HIP graph (launch cost removed)
    g = oa.MotionGraph(env, A, B, lvs)
    for _ in range(5):
        g.launch()
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(args.steps):
        g.launch()
    ev1.record(stream)
    torch.cuda.synchronize()
    graph_ms = ev0.elapsed_time(ev1) / args.steps
    graph_lat = []
    for _ in range(min(args.steps, 200)):
        s0 = time.perf_counter()
        g.launch()
        torch.cuda.synchronize()
        graph_lat.append((time.perf_counter() - s0) * 1e3)
    graph_same = bool(np.array_equal(g.out.cpu().numpy(), valid))
    g.close()
    # the two-stream pipeline captured: two independent batches on forked
    # streams in one graph, replayed K / 2 times (per-batch time reported)
    g2 = oa.MotionGraph.group(env, [(A, B), (A, B)], lvs)
    for _ in range(5):
        g2.launch()
    torch.cuda.synchronize()
    reps = max(1, args.steps // 2)
    ev0.record(stream)
    for _ in range(reps):
        g2.launch()
    ev1.record(stream)
    torch.cuda.synchronize()
    graph2_ms = ev0.elapsed_time(ev1) / (2 * reps)
    graph2_same = all(bool(np.array_equal(o.cpu().numpy(), valid)) for o in g2.out)
    g2.close()
    # graphs (each its own workspace and output) replayed round-robin on the
    # picked streams, as the eager pipeline issues its batches (time per batch)
    def graphs_on_streams(n_streams):
        use = (picked * n_streams)[:n_streams]
        gs = [oa.MotionGraph(env, A, B, lvs) for _ in use]
        handles = [s.cuda_stream for s in use]
        for _ in range(3):
            for g_, h in zip(gs, handles):
                g_.launch(h)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(use[0])
        for s in use[1:]:
            s.wait_event(e0)
        for i in range(args.steps):
            gs[i % len(gs)].launch(handles[i % len(gs)])
        for s in use[1:]:
            use[0].wait_stream(s)
        e1.record(use[0])
        torch.cuda.synchronize()
        same = all(bool(np.array_equal(g_.out.cpu().numpy(), valid)) for g_ in gs)
        for g_ in gs:
            g_.close()
        return e0.elapsed_time(e1) / args.steps, same

    graph_2s_ms, graph_2s_same = graphs_on_streams(2)
    graph_4s_ms, graph_4s_same = graphs_on_streams(len(picked))

    # ---- per-kernel durations: HIP events recorded by the library on the
    #      launch stream around every launch of the same K steps
    def kernel_times():
        oa.profile.reset()
        oa.profile.enable(True)
        for _ in range(args.steps):
            oa.check_motions(env, A, B, lvs, out=out)
        torch.cuda.synchronize()
        per = {}
        for name in ("motion_fused", "motion_endpoint", "motion_walk", "motion_chunk"):
            n, ms = oa.profile.query(name)
            if n:
                per[name] = {"launches": n, "avg_us": 1e3 * ms / n}
        oa.profile.enable(False)
        oa.profile.reset()
        return per

    kernels = kernel_times()

    # ---- algorithmic work in the reference order, for the roofline
    work = oa.motion_work(env, A, B, lvs)
    flops_step = oa.work_flops(work)
    w_end = oa.validity_work(env, B.float().contiguous())  # the s2 checks of every edge
    flops_end = oa.work_flops(w_end)
    share = {"motion_fused": flops_step, "motion_endpoint": flops_end, "motion_walk": flops_step - flops_end,
             "motion_chunk": flops_step - flops_end}
    if t["motion_strategy"] == 0:
        share["motion_walk"] = flops_step
    # the dominant kernel is the one with the most time per step; the chunk
    # walk's two launches (one event scope) count as one kernel (its work is
    # not split by phase)
    dom = max(kernels, key=lambda k: kernels[k]["avg_us"])
    dom_flops = share[dom]
    dom_us = kernels[dom]["avg_us"]
    achieved = dom_flops / (dom_us * 1e-6) / 1e12
    traffic, traffic_src = hbm_traffic(dom, t)
    # the s2 pass's figure beside it
    roof_s2 = None
    if dom != "motion_endpoint" and "motion_endpoint" in kernels:
        a2 = flops_end / (kernels["motion_endpoint"]["avg_us"] * 1e-6) / 1e12
        roof_s2 = {"kernel": "motion_endpoint", "achieved": a2, "frac": a2 / PEAK_FP32_TFLOPS,
                   "kernel_avg_us": kernels["motion_endpoint"]["avg_us"], "algorithmic_flops_per_launch": flops_end,
                   "traffic": hbm_traffic("motion_endpoint", t)[0]}

    # the rest (plain path, served stream, PRM, rows) with the one-stream shape
    oa.set_tuning(**t_one)
    # the same kernels with the one-stream shape (the shape for one batch at a time)
    roof_one = None
    if t_one != t:
        k1 = kernel_times()
        roof_one = {"tuning": t_one}
        for kname, fl in (("motion_chunk", flops_step - flops_end), ("motion_endpoint", flops_end)):
            if kname in k1:
                a1 = fl / (k1[kname]["avg_us"] * 1e-6) / 1e12
                roof_one[kname] = {"kernel_avg_us": k1[kname]["avg_us"], "achieved": a1,
                                   "frac": a1 / PEAK_FP32_TFLOPS, "traffic": hbm_traffic(kname, t_one)[0]}

    # ---- plain one-launch-per-call path on a prefix of the same edges
    P = min(args.plain_edges, N_EDGES)
    plain_lat = []
    plain_valid = np.zeros(P, np.uint8)
    for i in range(P):
        s0 = time.perf_counter()
        plain_valid[i] = oa.check_motion(env, a_rows[i], b_rows[i], lvs)
        plain_lat.append((time.perf_counter() - s0) * 1e3)
    plain_rate = P / (sum(plain_lat) / 1e3)

    # ---- CPU restatement on a bounded sample (repeated full passes) + parity
    cpu = None
    cpu_all = None
    parity = None
    if not args.no_cpu:
        from oracle import oracle_py

        oenv = oracle_py.OracleEnv(spheres, boxes)
        ref_v, _ = oracle_py.check_motions(oenv, a_rows, b_rows, lvs)
        parity = bool(np.array_equal(ref_v, valid)) and bool(np.array_equal(ref_v[:P], plain_valid))

        def cpu_rate(threads, seconds):
            done, t0 = 0, time.perf_counter()
            while time.perf_counter() - t0 < seconds:
                oracle_py.check_motions(oenv, a_rows, b_rows, lvs, threads=threads)
                done += N_EDGES
            return done, time.perf_counter() - t0

        host = host_cpu()
        nthr = args.cpu_threads or cpu_threads_default()
        done, cpu_s = cpu_rate(nthr, args.cpu_seconds)
        cpu = {"value": done / cpu_s, "unit": "edge-checks/s", "cores": nthr, "kind": "port",
               "sample": f"{done // N_EDGES} full passes over the 16384-edge batch ({cpu_s:.1f} s), {CPU_FLAGS}, "
                         f"{nthr} POSIX threads over edges (the host's CPU share of one GPU)",
               "host": host}
        done, cpu_s = cpu_rate(1, max(2.0, args.cpu_seconds / 3))
        cpu_all = {"value": done / cpu_s, "unit": "edge-checks/s", "cores": 1, "kind": "port",
                   "sample": f"{done // N_EDGES} full passes, {cpu_s:.1f} s, one thread"}

    stream_res = None
    if args.stream_requests > 0:
        stream_res = served_stream(oa, args.stream_requests, args.stream_qps, args.stream_envs, args.plain_workers,
                                   log_path=args.benchmark_log, run_plain=args.plain_workers > 0)
    if args.stream_requests > 0:
        # the same stream with RRT-Connect requests only (the plain path keeps up with the offered rate there)
        stream_res["rrt_connect_only"] = served_stream(oa, args.stream_requests, args.stream_qps, args.stream_envs,
                                                       args.plain_workers, run_plain=args.plain_workers > 0,
                                                       prm_every=0)
    if args.stream_requests > 0 and not args.no_stream_sweeps:
        # the same stream served alone at higher offered rates (saturation sweep)
        stream_res["served_qps_sweep"] = {
            str(int(q)): served_stream(oa, args.stream_requests, q, args.stream_envs, 0, run_plain=False)["served"]
            for q in (1000.0, 5000.0, 20000.0)}
        # saturated rates: the served path at its highest offered rate against
        # the plain path, which the 200 QPS stream already saturates (its
        # request rate falls below the offered one)
        plain = stream_res.get("plain")
        if plain is not None and plain["throughput_rps"] < 0.9 * args.stream_qps:
            top = max(v["edge_checks_per_s"] for v in stream_res["served_qps_sweep"].values())
            stream_res["saturated_edge_checks_served_over_plain"] = top / plain["edge_checks_per_s"]
        # server settings: obstacle partitions of the motion check (0 = by motion length) and workgroup size;
        # HIP events around each slice launch
        stream_res["served_config_sweep"] = {
            f"parts{parts}_t{threads}": served_stream(oa, args.stream_requests, args.stream_qps, args.stream_envs, 0,
                                                      check_parts=parts, block_threads=threads, profile_slices=True,
                                                      run_plain=False, prm_every=0)["served"]
            for parts, threads in ((0, 512), (1, 512), (2, 512), (4, 512), (9, 512), (0, 1024))}

    prm_res = None
    if args.prm_samples > 0:
        prm_res = prm_batch(oa, args.prm_samples, args.prm_cpu_samples,
                            args.cpu_threads or cpu_threads_default(), not args.no_cpu)

    rows_res = None
    if not args.no_rows:
        oa.set_tuning(**t)  # the batch's tuned shapes (the PRM leg tunes its own)
        rows_res = scope_rows(oa, not args.no_cpu, args.cpu_threads or cpu_threads_default(),
                              only=set(args.rows.split(",")) if args.rows else None,
                              sparse_tuning=json.loads(args.sparse_tuning) if args.sparse_tuning else None)

    res = {
        "metric": "edge-checks/s + p50/p99 plan latency, served vs plain GPU path, bit-exact",
        "value": value,
        "unit": "edge-checks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (FK/collision), fp64 (segment count, interpolation)",
        "data": "synthetic (seeded std::mt19937 workload)",
        "config": {"workload": "uneven motion validation: 7-DOF arm, 128 spheres + 128 boxes, 16384 edges "
                               "lognormal(-1,1) length, 0.01 rad, bisection + first-collision exit",
                   "edges": N_EDGES, "obstacles": N_SPHERES + N_BOXES, "tuning": t, "streams": len(picked)},
        "streams_asked": args.streams,
        "stream_pair_checks": pair_checks,
        "one_stream": {"edge_checks_per_s": N_EDGES * args.steps / (single_ms / 1e3),
                       "ms_per_step": single_ms / args.steps, "tuning": t_one},
        "two_streams": {"edge_checks_per_s": N_EDGES * args.steps / (pair_ms / 1e3),
                        "ms_per_step": pair_ms / args.steps},
        "streams_same_result": streams_same,
        "kernels": kernels,
        "valid_fraction": float(valid.mean()),
        "step_latency_ms": {"p50": percentile(step_ms, 50), "p99": percentile(step_ms, 99)},
        "shape_sweep": sweep,
        "graph": {"ms_per_step": graph_ms, "edge_checks_per_s": N_EDGES / (graph_ms * 1e-3),
                  "step_latency_ms": {"p50": percentile(graph_lat, 50), "p99": percentile(graph_lat, 99)},
                  "same_result": graph_same,
                  "two_streams": {"ms_per_step": graph_2s_ms, "edge_checks_per_s": N_EDGES / (graph_2s_ms * 1e-3),
                                  "same_result": graph_2s_same,
                                  "note": "two graphs replayed alternately on two streams; time per batch"},
                  "all_streams": {"streams": len(picked), "ms_per_step": graph_4s_ms,
                                  "edge_checks_per_s": N_EDGES / (graph_4s_ms * 1e-3), "same_result": graph_4s_same,
                                  "note": "one graph per stream, replayed round-robin on the streams the timed "
                                          "steps use; time per batch"},
                  "two_batch_group": {"ms_per_step": graph2_ms, "edge_checks_per_s": N_EDGES / (graph2_ms * 1e-3),
                                      "same_result": graph2_same,
                                      "note": "two independent batches on forked streams in one graph; "
                                              "time per batch"}},
        "plain": {"edge_checks_per_s": plain_rate, "edges": P,
                  "call_latency_ms": {"p50": percentile(plain_lat, 50), "p99": percentile(plain_lat, 99)}},
        "batched_over_plain": value / plain_rate,
        "work_per_step": work,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": dom, "kernel_avg_us": dom_us, "algorithmic_flops_per_launch": dom_flops,
                     "dispatches": 2 if dom == "motion_chunk" else 1,
                     "compute": "FP32 VALU (compute-bound; MI355X FP32 vector peak = FP32 matrix peak)"},
        "roofline_s2_pass": roof_s2,
        "roofline_one_stream_shape": roof_one,
        # the whole step: the batch's reference-order work over the timed step
        # (both kernels, the streams overlapping)
        "roofline_step": {"achieved": flops_step / (ms_per_step * 1e-3) / 1e12,
                          "frac": flops_step / (ms_per_step * 1e-3) / 1e12 / PEAK_FP32_TFLOPS,
                          "algorithmic_flops_per_step": flops_step, "ms_per_step": ms_per_step},
        "cpu_baseline": cpu,
        "cpu_one_thread": cpu_all,
        "autotuned_us_per_batch": tuned_us,
        "autotuned_streams_us_per_batch": streams_us,
        "parity_bit_exact": parity,
        "served_stream": stream_res,
        "prm_batch": prm_res,
        "rows": rows_res,
        "timed_region_wall_s": wall,
    }
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
