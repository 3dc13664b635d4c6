"""GPU tests of the planning server (continuous batching) and the plain
per-call planner: plans equal the CPU RRT-Connect restatement, each request's
result is the same alone and in the mix, and the failure paths (deadline,
admission, cancellation, tree capacity) are forced from the host."""
import math
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
MAXD = 0.2 * 2 * math.pi * math.sqrt(7)


def lvs_config():
    ext = 2 * math.pi * math.sqrt(7)
    return ext * (0.01 / ext)


def valid_pairs(gpu, oracle, oenv, n, seed):
    """Valid start/goal pairs drawn from seeded configurations."""
    q = gpu.workload.configs(seed, 4000)
    ok = [row for row in q if oracle.state_valid(oenv, row.astype(np.float32))]
    assert len(ok) >= 2 * n
    return [(ok[2 * i], ok[2 * i + 1]) for i in range(n)]


@pytest.fixture(scope="module")
def scene(gpu, oracle):
    s, b = gpu.workload.obstacles(1020, 12, 12)
    env = gpu.Env(s, b)
    oenv = oracle.OracleEnv(s, b)
    pairs = valid_pairs(gpu, oracle, oenv, 10, 77)
    return env, oenv, pairs


def reqs_for(gpu, env, pairs, max_iters=3000, period=None):
    """Request i plans pair i with seed 100 + (i mod period)."""
    from ompl_amd.serve import make_request

    period = period or len(pairs)
    return [make_request(env, a, b, seed=100 + i % period, lvs=lvs_config(), max_iters=max_iters)
            for i, (a, b) in enumerate(pairs)]


def test_plain_planner_matches_oracle(gpu, oracle, scene):
    from ompl_amd.serve import SOLVED, plan_plain

    env, oenv, pairs = scene
    for i, r in enumerate(reqs_for(gpu, env, pairs[:5])):
        res, path = plan_plain(r)
        ok, it, opath, nodes = oracle.rrt_connect(oenv, pairs[i][0], pairs[i][1], 100 + i, MAXD, lvs_config(),
                                                  max_iters=3000)
        print(f"plain request {i}: status {res.status} iters {res.iterations} oracle {ok} {it}", flush=True)
        assert (res.status == SOLVED) == ok
        assert res.iterations == it and res.tree_nodes == nodes
        assert np.array_equal(path, opath)


# motion-check obstacle partitions (0 = by motion length) and workgroup sizes
@pytest.mark.parametrize("parts,threads", [(0, 1024), (9, 1024), (2, 1024), (1, 512), (4, 512)])
def test_served_equals_oracle_alone_and_in_mix(gpu, oracle, scene, parts, threads):
    from ompl_amd.serve import SOLVED, Server

    env, oenv, pairs = scene
    reqs = reqs_for(gpu, env, pairs)
    print(f"served check partitions {parts}, {threads} threads", flush=True)
    srv = Server(hbm_budget_bytes=1 << 30, max_active=64, max_nodes=8192, check_obstacle_parts=parts,
                 block_threads=threads)
    alone = []
    for r in reqs:  # one at a time
        rid = srv.submit(r)
        res = srv.wait(rid, 60000)
        alone.append((res.status, res.iterations, res.tree_nodes, srv.path(rid), res.motion_checks))
        srv.release(rid)
    ids = [srv.submit(r) for r in reqs]  # all together
    mix = []
    for rid in ids:
        res = srv.wait(rid, 60000)
        mix.append((res.status, res.iterations, res.tree_nodes, srv.path(rid), res.motion_checks))
    for i, (a, m) in enumerate(zip(alone, mix)):
        assert a[0] == m[0] and a[1] == m[1] and a[2] == m[2] and np.array_equal(a[3], m[3]), i
        assert a[4] == m[4], i
        ok, it, opath, nodes = oracle.rrt_connect(oenv, pairs[i][0], pairs[i][1], 100 + i, MAXD, lvs_config(),
                                                  max_iters=3000)
        assert (a[0] == SOLVED) == ok and a[1] == it and a[2] == nodes
        assert np.array_equal(a[3], opath)
    st = srv.stats()
    assert st["active"] == 0 and st["queued"] == 0 and st["used_bytes"] == 0
    assert sum(m[0] == SOLVED for m in mix) >= 5
    srv.close()


def test_served_motion_checks_equal_plain(gpu, scene):
    """The served and plain planners make the same growTree steps, so they
    check the same number of motions."""
    from ompl_amd.serve import Server, plan_plain

    env, _, pairs = scene
    reqs = reqs_for(gpu, env, pairs[:4])
    srv = Server(hbm_budget_bytes=1 << 30, max_active=8, max_nodes=8192)
    ids = [srv.submit(r) for r in reqs]
    for r, rid in zip(reqs, ids):
        res = srv.wait(rid, 60000)
        pres, ppath = plan_plain(r)
        assert res.motion_checks == pres.motion_checks > 0
        assert res.iterations == pres.iterations and np.array_equal(srv.path(rid), ppath)
    srv.close()


def test_continuous_batching_more_requests_than_slots(gpu, oracle, scene):
    from ompl_amd.serve import SOLVED, FAILED, Server

    env, oenv, pairs = scene
    reqs = reqs_for(gpu, env, pairs * 3, period=len(pairs))
    srv = Server(hbm_budget_bytes=1 << 30, max_active=4, max_nodes=4096)
    ids = [srv.submit(r) for r in reqs]
    results = [srv.wait(rid, 120000) for rid in ids]
    assert all(r.status in (SOLVED, FAILED) for r in results)
    # identical requests (same pair and seed) give identical plans regardless of admission order
    for i in range(len(pairs)):
        a, b = ids[i], ids[i + len(pairs)]
        assert np.array_equal(srv.path(a), srv.path(b)) and results[i].iterations == results[i + len(pairs)].iterations
    srv.close()


def test_failure_paths_forced_from_host(gpu, scene):
    from ompl_amd.serve import CANCELLED, DEADLINE, REJECTED, TREE_FULL, Server, make_request

    env, _, pairs = scene
    a, b = pairs[0]
    lvs = lvs_config()
    # admission: a single request's footprint exceeds the whole budget
    srv = Server(hbm_budget_bytes=1000, max_active=4, max_nodes=1024)
    rid = srv.submit(make_request(env, a, b, 1, lvs))
    assert srv.wait(rid, 1000).status == REJECTED
    srv.close()
    srv = Server(hbm_budget_bytes=1 << 30, max_active=2, max_nodes=1024)
    # deadline already expired
    rid = srv.submit(make_request(env, a, b, 1, lvs, max_iters=10 ** 7, deadline_ms=1e-6))
    assert srv.wait(rid, 5000).status == DEADLINE
    # tree capacity: 3 nodes per tree cannot hold a plan
    rid = srv.submit(make_request(env, a, b, 2, lvs, max_iters=10 ** 6, max_nodes=3))
    assert srv.wait(rid, 5000).status == TREE_FULL
    # cancellation of a long request (a far goal in steps of 1e-4: thousands of
    # slices, so the cancellation is applied between two of them)
    rid = srv.submit(make_request(env, a, b, 3, lvs, max_iters=10 ** 7, rng=1e-4, max_nodes=1024))
    srv.cancel(rid)
    assert srv.wait(rid, 5000).status == CANCELLED
    assert srv.stats()["used_bytes"] == 0
    srv.close()


def test_server_rejects_unbuilt_check_shapes(gpu):
    from ompl_amd._lib import OmplAmdError
    from ompl_amd.serve import Server

    for lanes, parts in ((8, 2), (8, 8), (7, 8), (6, 0), (7, 3)):
        with pytest.raises(OmplAmdError):
            Server(hbm_budget_bytes=1 << 30, max_active=2, max_nodes=64, check_obstacle_parts=parts,
                   check_lanes_per_state=lanes)
    with pytest.raises(OmplAmdError):
        Server(hbm_budget_bytes=1 << 30, max_active=2, max_nodes=64, block_threads=256)


def test_admission_bound_by_hbm_budget(gpu, scene):
    """With a pool of 2.5 request regions at most two requests run at once: the
    third waits in the queue for memory, not for a slot, and its plan is the
    one it makes alone."""
    from ompl_amd.serve import SOLVED, Server, footprint, make_request

    env, _, pairs = scene
    lvs = lvs_config()
    fp = footprint(2048)
    reqs = [make_request(env, pairs[i][0], pairs[i][1], 500 + i, lvs, max_iters=3000) for i in range(6)]
    alone = Server(hbm_budget_bytes=1 << 28, max_active=8, max_nodes=2048)
    ref = []
    for r in reqs:
        rid = alone.submit(r)
        res = alone.wait(rid, 60000)
        ref.append((res.status, res.iterations, alone.path(rid)))
    alone.close()
    srv = Server(hbm_budget_bytes=int(2.5 * fp), max_active=8, max_nodes=2048)
    ids = [srv.submit(r) for r in reqs]
    peak = 0
    while True:
        st = srv.stats()
        peak = max(peak, st["active"])
        assert st["used_bytes"] <= st["budget_bytes"]
        if st["active"] == 0 and st["queued"] == 0:
            break
        time.sleep(2e-4)
    res = [srv.wait(rid, 60000) for rid in ids]
    assert peak <= 2
    assert any(r.queued_ms > 0.0 for r in res[2:])
    for r, rid, (st0, it0, p0) in zip(res, ids, ref):
        assert r.status == st0 and r.iterations == it0 and np.array_equal(srv.path(rid), p0)
    assert sum(r.status == SOLVED for r in res) >= 3
    srv.close()


def test_release_right_after_finish_is_safe(gpu, scene):
    """Requests released as soon as they report a final status: the server
    publishes status, path and timing together, so nothing of a released
    request is touched afterwards (repeated, with more requests than slots)."""
    from ompl_amd.serve import SOLVED, Server, make_request

    env, _, pairs = scene
    lvs = lvs_config()
    srv = Server(hbm_budget_bytes=1 << 28, max_active=3, max_nodes=2048)
    for rep in range(3):
        ids = [srv.submit(make_request(env, a, b, 900 + i, lvs, max_iters=3000)) for i, (a, b) in enumerate(pairs)]
        left = set(ids)
        while left:
            for rid in list(left):
                r = srv.poll(rid)
                if r.status >= SOLVED:
                    assert r.latency_ms >= r.run_ms >= 0.0 and r.latency_ms >= r.queued_ms
                    srv.release(rid)
                    left.discard(rid)
            time.sleep(1e-4)
    st = srv.stats()
    assert st["active"] == 0 and st["used_bytes"] == 0
    srv.close()


def stream_requests(gpu, oracle, n_requests=64, seed=8, n_envs=32):
    """The served stream's workload (config 5): environments with log-uniform
    16-1,024 obstacles (half spheres, half boxes of the motion benchmark's
    sizes; the densest leave no collision-free configuration and drop out),
    plus two environments of 1,024 small obstacles (radii and half-extents
    U[0.01, 0.03]) so the largest obstacle count is served; start and goal
    are valid seeded configurations; max_iters 2,000 as in the benchmark."""
    import torch

    from ompl_amd.serve import make_request

    rng = np.random.default_rng(seed)
    scenes = []
    for e in range(n_envs):
        n_obs = int(round(math.exp(rng.uniform(math.log(16), math.log(1024)))))
        scenes.append((1000 + e, n_obs - n_obs // 2, n_obs // 2, (0.05, 0.15), (0.05, 0.2)))
    scenes += [(1100, 512, 512, (0.01, 0.03), (0.01, 0.03)), (1101, 512, 512, (0.01, 0.03), (0.01, 0.03))]
    envs = []
    for env_seed, ns, nb, r, h in scenes:
        s, b = gpu.workload.obstacles(env_seed, ns, nb, r=r, h=h)
        env = gpu.Env(s, b)
        q = gpu.workload.configs(env_seed + 1000, 4096)
        v = gpu.validate_states(env, torch.from_numpy(q.T.copy()).cuda()).cpu().numpy()
        ok = q[v == 1]
        if len(ok) >= 2:
            envs.append((env, oracle.OracleEnv(s, b), ok, ns + nb))
    reqs = []
    for i in range(n_requests + 4):
        k = i % (len(envs) - 2) if i < n_requests else len(envs) - 2 + (i % 2)
        env, oenv, ok, n_obs = envs[k]
        a, b = ok[rng.integers(len(ok))], ok[rng.integers(len(ok))]
        reqs.append((make_request(env, a, b, seed=10_000 + i, lvs=lvs_config(), max_iters=2000), oenv, a, b,
                     10_000 + i, n_obs))
    return reqs, envs  # the caller keeps the environments alive while requests use them


def test_served_stream_32_environments(gpu, oracle):
    """Config 5's stream at its shape: 68 requests over 27 environments of
    16-1,024 obstacles; each request's plan is the same alone and in the mix
    (all submitted at once, more than the server runs at a time) and equal to
    the oracle's RRT-Connect."""
    import concurrent.futures as cf

    from ompl_amd.serve import Server

    reqs, envs = stream_requests(gpu, oracle)
    assert max(r[5] for r in reqs) == 1024 and min(r[5] for r in reqs) <= 32
    srv = Server(hbm_budget_bytes=1 << 30, max_active=24, max_nodes=4096)
    alone = []
    for r, *_ in reqs:
        rid = srv.submit(r)
        res = srv.wait(rid, 60000)
        alone.append((res.status, res.iterations, res.tree_nodes, res.motion_checks, srv.path(rid)))
        srv.release(rid)
    print(f"stream: {len(reqs)} requests served alone", flush=True)
    ids = [srv.submit(r) for r, *_ in reqs]
    mix = []
    for rid in ids:
        res = srv.wait(rid, 120000)
        mix.append((res.status, res.iterations, res.tree_nodes, res.motion_checks, srv.path(rid)))
    st = srv.stats()
    srv.close()
    assert st["active"] == 0 and st["used_bytes"] == 0
    del envs

    def plan(item):
        _, oenv, a, b, seed, _ = item
        return oracle.rrt_connect(oenv, a, b, seed, MAXD, lvs_config(), max_iters=2000)

    with cf.ThreadPoolExecutor(16) as ex:
        ref = list(ex.map(plan, reqs))
    from ompl_amd.serve import SOLVED

    for i, (a_, m_, (ok, it, opath, nodes)) in enumerate(zip(alone, mix, ref)):
        assert a_[:4] == m_[:4] and np.array_equal(a_[4], m_[4]), i
        assert (a_[0] == SOLVED) == ok and a_[1] == it and a_[2] == nodes, i
        assert np.array_equal(a_[4], opath), i
    assert sum(a_[0] == SOLVED for a_ in alone) >= 20


def test_environment_in_use_cannot_be_destroyed(gpu, scene):
    """An environment that a queued or running request reads cannot be
    destroyed (the call fails and frees nothing); once the request finishes,
    it can."""
    from ompl_amd._lib import OmplAmdError
    from ompl_amd.serve import Server, make_request

    _, _, pairs = scene
    s, b = gpu.workload.obstacles(1020, 12, 12)
    env = gpu.Env(s, b)
    srv = Server(hbm_budget_bytes=1 << 28, max_active=2, max_nodes=2048)
    # a long request (range 1e-4 rad between configurations radians apart), cancelled afterwards
    rid = srv.submit(make_request(env, pairs[0][0], pairs[0][1], 5, lvs_config(), max_iters=10 ** 7, rng=1e-4))
    with pytest.raises(OmplAmdError, match="in use"):
        env.close()
    assert env.handle is not None and env.handle.value
    srv.cancel(rid)
    assert srv.wait(rid, 60000).status >= 2
    env.close()
    assert env.handle is None
    srv.close()


def test_served_prm_requests_equal_oracle(gpu, oracle, scene):
    """PRM requests: start and goal as milestones 0 and 1 of a batch roadmap,
    shortest path over its valid edges.  Served path, milestone count and
    checked candidate edges equal the oracle's; the plain host PRM (one launch
    per check) gives the same; an invalid goal fails."""
    from ompl_amd.serve import FAILED, PRM, SOLVED, Server, make_request, plan_plain

    env, oenv, pairs = scene
    lvs = lvs_config()
    reqs = [make_request(env, a, b, 700 + i, lvs, planner=PRM, prm_samples=1500, prm_k=10)
            for i, (a, b) in enumerate(pairs[:6])]
    bad = gpu.workload.configs(5, 4000)
    bad = next(q for q in bad if not oracle.state_valid(oenv, q.astype(np.float32)))
    reqs.append(make_request(env, pairs[0][0], bad, 800, lvs, planner=PRM, prm_samples=500, prm_k=10))
    srv = Server(hbm_budget_bytes=1 << 30, max_active=8, max_nodes=2048)
    ids = [srv.submit(r) for r in reqs]
    res = [srv.wait(rid, 60000) for rid in ids]
    paths = [srv.path(rid) for rid in ids]
    srv.close()
    solved = 0
    for i, (r, p) in enumerate(zip(res, paths)):
        goal = pairs[i][1] if i < 6 else bad
        ok, opath, nv, ne = oracle.prm_plan(oenv, pairs[min(i, 5) if i < 6 else 0][0], goal, 700 + i if i < 6 else 800,
                                            1500 if i < 6 else 500, 10, lvs, threads=8)
        print(f"prm request {i}: status {r.status} milestones {r.tree_nodes} oracle {ok} {nv}", flush=True)
        assert (r.status == SOLVED) == ok and r.iterations == 1
        assert r.tree_nodes == nv and r.motion_checks == ne
        assert np.array_equal(p, opath)
        solved += ok
    assert res[-1].status == FAILED and solved >= 3
    for i in (0, 1):
        pres, ppath = plan_plain(reqs[i])
        assert pres.status == res[i].status and pres.tree_nodes == res[i].tree_nodes
        assert pres.motion_checks == res[i].motion_checks and np.array_equal(ppath, paths[i])


def test_rrt_connect_and_prm_requests_alone_and_in_mix(gpu, scene):
    """RRT-Connect and PRM requests served together (PRM on the server's second
    stream beside the RRT-Connect slices) give the results each gives alone."""
    from ompl_amd.serve import PRM, Server, make_request

    env, _, pairs = scene
    lvs = lvs_config()
    reqs = []
    for i, (a, b) in enumerate(pairs):
        reqs.append(make_request(env, a, b, 900 + i, lvs, max_iters=3000,
                                 planner=PRM if i % 3 == 0 else 0, prm_samples=1200))
    srv = Server(hbm_budget_bytes=1 << 30, max_active=8, max_nodes=4096)
    alone = []
    for r in reqs:
        rid = srv.submit(r)
        x = srv.wait(rid, 60000)
        alone.append((x.status, x.iterations, x.tree_nodes, x.motion_checks, srv.path(rid)))
    ids = [srv.submit(r) for r in reqs]
    for i, rid in enumerate(ids):
        x = srv.wait(rid, 60000)
        m = (x.status, x.iterations, x.tree_nodes, x.motion_checks, srv.path(rid))
        assert alone[i][:4] == m[:4] and np.array_equal(alone[i][4], m[4]), i
    st = srv.stats()
    assert st["active"] == 0 and st["queued"] == 0 and st["used_bytes"] == 0
    srv.close()


@pytest.mark.parametrize("intermediate", [False, True])
def test_start_goal_lists_and_intermediate_states(gpu, oracle, intermediate):
    """Requests listing several start and goal states (some out of bounds or in
    collision) with and without addIntermediateStates: served (alone and in
    the mix) and plain plans equal the oracle's status, iterations, tree sizes
    and paths; the goal schedule puts later goal states into the goal tree."""
    from ompl_amd.serve import Server, make_request, plan_plain

    s, b = gpu.workload.obstacles(1003, 24, 24)
    env = gpu.Env(s, b)
    oenv = oracle.OracleEnv(s, b)
    q = gpu.workload.configs(1003 + 77, 3000)
    ok = [row for row in q if oracle.state_valid(oenv, row.astype(np.float32))]
    bad = [row for row in gpu.workload.configs(1003 + 91, 400) if not oracle.state_valid(oenv, row.astype(np.float32))]
    outside = np.array([3.2, 0, 0, 0, 0, 0, 0], np.float64)
    cases = []
    for i in range(6):
        starts = [bad[i], outside, ok[10 * i], ok[10 * i + 1]]
        goals = [bad[i + 1], ok[10 * i + 2], bad[i + 2], ok[10 * i + 3], ok[10 * i + 4]]
        cases.append((starts, goals))
    cases.append(([bad[0], outside], [ok[0]]))           # no usable start
    cases.append(([ok[1]], [bad[1], bad[2], outside]))   # no usable goal
    reqs = [make_request(env, None, None, seed=60 + i, lvs=lvs_config(), max_iters=300, rng=1.0, starts=st,
                         goals=gl, intermediate_states=intermediate) for i, (st, gl) in enumerate(cases)]
    srv = Server(hbm_budget_bytes=1 << 30, max_active=16, max_nodes=1 << 15)
    alone = []
    for r in reqs:
        rid = srv.submit(r)
        res = srv.wait(rid, 60000)
        alone.append((res.status, res.iterations, res.tree_nodes, srv.path(rid)))
        srv.release(rid)
    ids = [srv.submit(r) for r in reqs]
    mix = [(lambda res, rid: (res.status, res.iterations, res.tree_nodes, srv.path(rid)))(srv.wait(rid, 60000), rid)
           for rid in ids]
    names = {0: 3, 1: 2, 2: 8, 3: 9}  # oracle status -> server status
    solved = 0
    for i, (st, gl) in enumerate(cases):
        ost, oit, opath, onodes = oracle.rrt_connect_ex(oenv, st, gl, 60 + i, 1.0, lvs_config(), max_iters=300,
                                                       intermediate=intermediate)
        print(f"case {i}: served {alone[i][:3]} oracle {(ost, oit, onodes)}", flush=True)
        for got in (alone[i], mix[i]):
            assert got[0] == names[ost] and got[1] == oit and got[2] == onodes, (i, got[:3], ost, oit, onodes)
            assert np.array_equal(got[3], opath), i
        pres, ppath = plan_plain(reqs[i], cap=1 << 15)
        assert (pres.status, pres.iterations, pres.tree_nodes) == (names[ost], oit, onodes), i
        assert np.array_equal(ppath, opath), i
        solved += ost == 1
    assert solved >= 3
    srv.close()
