"""GPU parity of state validity and motion validation against the CPU oracle.
Bit-exact: validity and motion booleans, checked-state counts and lastValid
fractions must equal the oracle's on the same seeded inputs."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
PI = math.pi
LVS = 0.01  # 0.01 rad resolution: 2*pi*sqrt(7) * (0.01 / (2*pi*sqrt(7)))


def lvs_config():
    ext = 2 * PI * math.sqrt(7)
    return ext * (0.01 / ext)


def to_dev(x, dtype=None):
    import torch

    return torch.from_numpy(np.ascontiguousarray(x)).to("cuda") if dtype is None else \
        torch.from_numpy(np.ascontiguousarray(x)).to("cuda", dtype)


def env_pair(gpu, oracle, seed, ns, nb):
    s, b = gpu.workload.obstacles(seed, ns, nb)
    return gpu.Env(s, b), oracle.OracleEnv(s, b), (s, b)


def test_fk_bit_exact(gpu, oracle):
    rng = np.random.default_rng(0)
    q = rng.uniform(-PI, PI, (2000, 7)).astype(np.float32)
    q[:10] = 0.0
    q[10:20] = rng.uniform(-20, 20, (10, 7))
    got = gpu.arm_spheres(to_dev(q.T.copy())).cpu().numpy()
    for i in range(q.shape[0]):
        assert np.array_equal(got[i], oracle.arm_spheres(q[i])), i


# (7, x): link layout with bounding-sphere culling, 8 / 4 / 2 / 1 states per wave
SHAPES = [(1, 1), (2, 1), (4, 1), (8, 1), (8, 2), (8, 4), (8, 8), (7, 1), (7, 2), (7, 4), (7, 9)]
# (lanes, parts, shared FK)
VSHAPES = [(lps, ops, 0) for lps, ops in SHAPES] + [(8, 2, 1), (8, 4, 1), (8, 8, 1)]


@pytest.mark.parametrize("lps,ops,shared", VSHAPES)
@pytest.mark.parametrize("ns,nb", [(64, 0), (128, 128), (0, 40), (0, 0), (3, 1)])
def test_validate_states_bit_exact(gpu, oracle, lps, ops, shared, ns, nb):
    print(f"validate shape=({lps},{ops}) shared={shared} ns={ns} nb={nb}", flush=True)
    gpu.set_tuning(validity_lanes_per_state=lps, validity_obstacle_parts=ops, validity_shared_fk=shared)
    env, oenv, _ = env_pair(gpu, oracle, 1000 + ns + nb, ns, nb)
    q = gpu.workload.configs(1, 65536 if (ns, nb) == (64, 0) else 8192)
    qf = q.astype(np.float32).T.copy()
    got = gpu.validate_states(env, to_dev(qf)).cpu().numpy()
    ref = oracle.validate_states(oenv, qf)
    assert np.array_equal(got, ref)
    # FP64 input: cast per joint on device
    got64 = gpu.validate_states(env, to_dev(q.T.copy())).cpu().numpy()
    assert np.array_equal(got64, ref)
    if ns + nb == 0:
        assert ref.all()
    gpu.set_tuning(validity_lanes_per_state=8, validity_obstacle_parts=2, validity_shared_fk=0)


@pytest.mark.parametrize("n", [1, 63, 65, 129, 1000])
def test_validate_ragged_sizes(gpu, oracle, n):
    env, oenv, _ = env_pair(gpu, oracle, 7, 32, 32)
    q = gpu.workload.configs(n, n).astype(np.float32).T.copy()
    for lps, ops, shared in ((1, 1, 0), (8, 1, 0), (8, 8, 0), (8, 8, 1), (8, 2, 1), (7, 9, 0), (7, 1, 0), (7, 4, 0)):
        gpu.set_tuning(validity_lanes_per_state=lps, validity_obstacle_parts=ops, validity_shared_fk=shared)
        assert np.array_equal(gpu.validate_states(env, to_dev(q)).cpu().numpy(), oracle.validate_states(oenv, q))
    gpu.set_tuning(validity_lanes_per_state=8, validity_obstacle_parts=2, validity_shared_fk=0)


def box_distance(p, lo, hi):
    return float(np.linalg.norm(p - np.clip(p, lo, hi)))


def near_contact_envs(oracle, n_envs, seed):
    """Environments of one sphere and one box obstacle, each within ~1e-6 m of
    touching one robot sphere of a base state q0 and clear (by > 1 mm) of the
    other 27 (sphere obstacle: centre at distance R + r from the robot sphere;
    box: a face at distance r from it), with 64 states q0 + 1e-6 noise per
    environment, so that the exact FP32 tests fall on both sides of contact."""
    rng = np.random.default_rng(seed)
    made = 0
    while made < n_envs:
        q0 = rng.uniform(-PI, PI, 7)
        sph = oracle.arm_spheres(q0.astype(np.float32)).astype(np.float64)
        j, k = rng.integers(0, 28, 2)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        R = rng.uniform(0.01, 0.08)
        o = sph[j] + d * (R + 0.05) * (1 + rng.uniform(-3e-7, 3e-7))
        others = np.delete(np.arange(28), j)
        if np.min(np.linalg.norm(sph[others] - o, axis=1)) < R + 0.05 + 1e-3:
            continue
        h = rng.uniform(0.01, 0.06, 3)
        axis = rng.integers(0, 3)
        c = sph[k].copy()
        c[axis] += rng.choice([-1.0, 1.0]) * (h[axis] + 0.05 * (1 + rng.uniform(-3e-6, 3e-6)))
        lo, hi = c - h, c + h
        if min(box_distance(sph[i], lo, hi) for i in range(28) if i != k) < 0.05 + 1e-3:
            continue
        made += 1
        spheres = np.array([[*o, R]], np.float32)
        boxes = np.array([[*c, *h]], np.float32)
        q = (q0[None, :] + rng.normal(size=(64, 7)) * 1e-6).astype(np.float32)
        q[0] = q0
        yield spheres, boxes, q


@pytest.mark.parametrize("lps,ops", [(7, 9), (7, 4), (7, 2), (7, 1), (8, 8), (1, 1)])
def test_validate_near_contact(gpu, oracle, lps, ops):
    """Obstacles within ~1e-6 m of contact: every shape, the link layout with
    its bound culling included, reproduces the oracle's exact FP32 decisions."""
    gpu.set_tuning(validity_lanes_per_state=lps, validity_obstacle_parts=ops, validity_shared_fk=0)
    n_valid = n_total = 0
    for spheres, boxes, q in near_contact_envs(oracle, 60, 31):
        env = gpu.Env(spheres, boxes)
        oenv = oracle.OracleEnv(spheres, boxes)
        qf = np.ascontiguousarray(q.T)
        got = gpu.validate_states(env, to_dev(qf)).cpu().numpy()
        ref = oracle.validate_states(oenv, qf)
        assert np.array_equal(got, ref)
        n_valid += int(ref.sum())
        n_total += ref.size
        env.close()
    assert 0 < n_valid < n_total  # contact decided both ways
    gpu.set_tuning(validity_lanes_per_state=8, validity_obstacle_parts=2, validity_shared_fk=0)


def test_link_layout_large_env(gpu, oracle):
    """An environment whose link-layout staging exceeds 64 KiB of LDS (4,000
    spheres + 48 boxes) is checked with the (8, 8) shape instead: same results."""
    gpu.set_tuning(validity_lanes_per_state=7, validity_obstacle_parts=9, motion_lanes_per_state=7,
                   motion_obstacle_parts=9, motion_strategy=3)
    env, oenv, _ = env_pair(gpu, oracle, 1005, 4000, 48)
    q = gpu.workload.configs(3, 2048).astype(np.float32).T.copy()
    ref = oracle.validate_states(oenv, q)
    assert np.array_equal(gpu.validate_states(env, to_dev(q)).cpu().numpy(), ref)
    a, b = motion_case(gpu, 6, 512)
    lvs = lvs_config()
    ref_v, ref_c = oracle.check_motions(oenv, a, b, lvs)
    v, c = gpu.check_motions(env, to_dev(a.T.copy()), to_dev(b.T.copy()), lvs, exact_count=True)
    assert np.array_equal(v.cpu().numpy(), ref_v) and np.array_equal(c.cpu().numpy(), ref_c)
    gpu.set_tuning(**DEFAULT)


def test_env_rejects_non_finite_obstacles(gpu):
    from ompl_amd._lib import OmplAmdError

    good_s, good_b = np.array([[0.5, 0.5, 0.5, 0.1]], np.float32), np.array([[0.5, 0, 0, 0.1, 0.1, 0.1]], np.float32)
    for bad in (np.nan, np.inf, -np.inf):
        for k in range(4):
            s = good_s.copy()
            s[0, k] = bad
            with pytest.raises(OmplAmdError):
                gpu.Env(s, good_b)
        for k in range(6):
            b = good_b.copy()
            b[0, k] = bad
            with pytest.raises(OmplAmdError):
                gpu.Env(good_s, b)
    gpu.Env(good_s, good_b).close()


@pytest.mark.parametrize("eshape,wshape", [((7, 4), (7, 9)), ((8, 2), (8, 1)), ((7, 1), (7, 2))])
def test_results_independent_of_obstacle_order(gpu, oracle, eshape, wshape):
    """The environment keeps its own staging order of the obstacles; validity and
    motion results must not depend on the order the caller lists them in."""
    s, b = gpu.workload.obstacles(1001, 128, 128)
    rng = np.random.default_rng(5)
    envs = [gpu.Env(s, b), gpu.Env(s[::-1].copy(), b[::-1].copy()),
            gpu.Env(s[rng.permutation(len(s))], b[rng.permutation(len(b))])]
    q = gpu.workload.configs(12, 4096).astype(np.float32)
    a, bb = motion_case(gpu, 13, 2048)
    A, B = to_dev(a.T.copy()), to_dev(bb.T.copy())
    set_strategy(gpu, 3, eshape, wshape)
    try:
        got = [(gpu.validate_states(e, to_dev(q.T.copy())).cpu().numpy(),
                gpu.check_motions(e, A, B, lvs_config()).cpu().numpy()) for e in envs]
    finally:
        gpu.set_tuning(**DEFAULT)
        for e in envs:
            e.close()
    oenv = oracle.OracleEnv(s, b)
    assert np.array_equal(got[0][0], oracle.validate_states(oenv, q.T.copy()))
    for v, m in got[1:]:
        assert np.array_equal(v, got[0][0])
        assert np.array_equal(m, got[0][1])


def reference_order_pairs(oracle, s, b, q):
    """Pairs a scalar checker visits for each state in the caller's obstacle
    order: robot spheres in order, each against the sphere obstacles then the
    box obstacles, stopping at the first hit (float32, the checker's formulas)."""
    f = np.float32
    rr = ((s[:, 3] + f(0.05)) * (s[:, 3] + f(0.05))).astype(f)
    lo, hi = (b[:, :3] - b[:, 3:]).astype(f), (b[:, :3] + b[:, 3:]).astype(f)
    sp = bp = 0
    for i in range(q.shape[0]):
        c = oracle.arm_spheres(q[i]).astype(f)  # 28 x 3
        d = s[None, :, :3] - c[:, None, :]
        hs = ((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]) < rr[None]
        e = c[:, None, :] - np.minimum(np.maximum(c[:, None, :], lo[None]), hi[None])
        hb = ((e[..., 0] * e[..., 0] + e[..., 1] * e[..., 1]) + e[..., 2] * e[..., 2]) < f(0.05) * f(0.05)
        seq = np.concatenate([hs, hb], 1).reshape(-1)  # robot-sphere major
        kind = np.tile(np.r_[np.zeros(len(s), bool), np.ones(len(b), bool)], c.shape[0])
        n = int(np.argmax(seq)) + 1 if seq.any() else seq.size
        bp += int(kind[:n].sum())
        sp += n - int(kind[:n].sum())
    return sp, bp


def test_work_accounting_in_caller_order(gpu, oracle):
    """The roofline's work count follows the caller's obstacle order, not the
    environment's staging order: listing the obstacles differently changes the
    count exactly as a scalar checker's would change."""
    s, b = gpu.workload.obstacles(1001, 128, 128)
    q = gpu.workload.configs(14, 300).astype(np.float32)
    for ss, bb in ((s, b), (s[::-1].copy(), b[::-1].copy())):
        env = gpu.Env(ss, bb)
        try:
            w = gpu.validity_work(env, to_dev(q.T.copy()))
        finally:
            env.close()
        sp, bp = reference_order_pairs(oracle, ss, bb, q)
        assert (w["states"], w["sphere_pairs"], w["box_pairs"]) == (q.shape[0], sp, bp)


def test_is_valid_single_call(gpu, oracle):
    env, oenv, _ = env_pair(gpu, oracle, 1000, 64, 0)
    q = gpu.workload.configs(5, 200)
    for i in range(q.shape[0]):
        assert gpu.is_valid(env, q[i]) == oracle.state_valid(oenv, q[i].astype(np.float32))


def motion_case(gpu, seed, n, mu=-1.0, sigma=1.0, lo=0.05, hi=5.0):
    a, b = gpu.workload.edges(seed, n, mu, sigma, lo, hi)
    return a, b


# (strategy, endpoint shape, walk shape)
STRATEGIES = [(3, (8, 2), (8, 1)), (3, (8, 8), (8, 8)), (3, (1, 1), (1, 1)), (3, (4, 1), (8, 4)),
              (1, (8, 1), (1, 1)), (1, (4, 1), (8, 2)), (0, (8, 1), (1, 1)), (0, (8, 1), (8, 8)),
              (2, (8, 1), (1, 1)), (2, (8, 2), (8, 8)), (2, (8, 4), (8, 1)),
              (4, (8, 8), (8, 8)), (4, (8, 2), (8, 4)), (4, (8, 4), (8, 1)), (4, (1, 1), (8, 2)),
              (3, (7, 9), (7, 9)), (3, (8, 8), (7, 9)), (1, (7, 9), (7, 9)), (0, (8, 1), (7, 9)),
              (4, (7, 9), (8, 8)), (3, (7, 4), (7, 2)), (3, (7, 1), (7, 4)), (1, (7, 2), (7, 1)),
              (0, (8, 1), (7, 2))]
DEFAULT = dict(motion_strategy=3, validity_lanes_per_state=8, validity_obstacle_parts=2,
               motion_lanes_per_state=8, motion_obstacle_parts=1)


def set_strategy(gpu, strategy, e, w):
    gpu.set_tuning(motion_strategy=strategy, validity_lanes_per_state=e[0], validity_obstacle_parts=e[1],
                   motion_lanes_per_state=w[0], motion_obstacle_parts=w[1])


@pytest.mark.parametrize("strategy,eshape,wshape", STRATEGIES)
def test_check_motions_config2_bit_exact(gpu, oracle, strategy, eshape, wshape):
    """Config: 16,384 edges, 128 spheres + 128 boxes, 0.01 rad resolution."""
    print(f"motions strategy={strategy} shapes={eshape},{wshape}", flush=True)
    set_strategy(gpu, strategy, eshape, wshape)
    env, oenv, _ = env_pair(gpu, oracle, 1001, 128, 128)
    a, b = motion_case(gpu, 4, 16384)
    lvs = lvs_config()
    ref_v, ref_c = oracle.check_motions(oenv, a, b, lvs)
    A, B = to_dev(a.T.copy()), to_dev(b.T.copy())
    v, c = gpu.check_motions(env, A, B, lvs, exact_count=True)
    assert np.array_equal(v.cpu().numpy(), ref_v)
    assert np.array_equal(c.cpu().numpy(), ref_c)
    v2 = gpu.check_motions(env, A, B, lvs)  # no count: early exit at any hit
    assert np.array_equal(v2.cpu().numpy(), ref_v)
    gpu.set_tuning(**DEFAULT)


@pytest.mark.parametrize("strategy,eshape,wshape", STRATEGIES)
def test_check_motions_sparse_env_deep_walks(gpu, oracle, strategy, eshape, wshape):
    """Few obstacles: most edges are valid, so every bisection position runs."""
    set_strategy(gpu, strategy, eshape, wshape)
    env, oenv, _ = env_pair(gpu, oracle, 1002, 6, 4)
    a, b = motion_case(gpu, 9, 3000, mu=-0.5, sigma=1.0, lo=0.05, hi=5.0)
    lvs = lvs_config()
    ref_v, ref_c = oracle.check_motions(oenv, a, b, lvs)
    assert 0.2 < ref_v.mean() < 0.95
    v, c = gpu.check_motions(env, to_dev(a.T.copy()), to_dev(b.T.copy()), lvs, exact_count=True)
    assert np.array_equal(v.cpu().numpy(), ref_v)
    assert np.array_equal(c.cpu().numpy(), ref_c)
    gpu.set_tuning(**DEFAULT)


@pytest.mark.parametrize("strategy,eshape,wshape,wpc", [(3, (8, 2), (8, 1), 8), (3, (4, 1), (7, 1), 32),
                                                         (3, (7, 4), (7, 9), 8), (4, (8, 2), (8, 4), 16),
                                                         (1, (8, 1), (1, 1), 16), (2, (8, 1), (8, 1), 8),
                                                         (3, (7, 4), (7, 4), 2), (3, (7, 2), (7, 9), 4)])
def test_check_motions_many_survivors(gpu, oracle, strategy, eshape, wshape, wpc):
    """16,384 edges in a 16 + 16 obstacle scene: about 10,000 edges pass s2,
    more than the chunk walk numbers edge-major, so its chunk-major form walks
    every position; the s2 pass appends that many survivors."""
    print(f"many survivors strategy={strategy} shapes={eshape},{wshape} wpc={wpc}", flush=True)
    set_strategy(gpu, strategy, eshape, wshape)
    gpu.set_tuning(motion_waves_per_cu=wpc)
    env, oenv, _ = env_pair(gpu, oracle, 1002, 16, 16)
    a, b = motion_case(gpu, 9, 16384, mu=-0.5, sigma=1.0, lo=0.05, hi=5.0)
    lvs = lvs_config()
    ref_v, ref_c = oracle.check_motions(oenv, a, b, lvs, threads=8)
    assert ref_v.sum() > 4096
    A, B = to_dev(a.T.copy()), to_dev(b.T.copy())
    v, c = gpu.check_motions(env, A, B, lvs, exact_count=True)
    assert np.array_equal(v.cpu().numpy(), ref_v)
    assert np.array_equal(c.cpu().numpy(), ref_c)
    for _ in range(2):  # repeated batches reuse the counter sets
        assert np.array_equal(gpu.check_motions(env, A, B, lvs).cpu().numpy(), ref_v)
    gpu.set_tuning(**DEFAULT, motion_waves_per_cu=8)


@pytest.mark.parametrize("vblock,mblock,wpc,wshape", [(64, 64, 32, (7, 1)), (256, 1024, 32, (7, 1)),
                                                      (1024, 64, 32, (7, 1)), (256, 256, 32, (8, 1)),
                                                      (256, 256, 16, (7, 2))])
def test_edge_major_walk_with_more_waves_than_resident(gpu, oracle, vblock, mblock, wpc, wshape):
    """About 3,000 survivors (the walk's edge-major second phase) over a grid
    with more waves than the CUs hold at once: blocks of phase 2 that start
    late, after others have found invalid positions, must deal the same chunks
    as the early ones (no chunk skipped)."""
    print(f"edge-major walk blocks {vblock}/{mblock} wpc {wpc} shape {wshape}", flush=True)
    gpu.set_tuning(motion_strategy=3, validity_lanes_per_state=4, validity_obstacle_parts=1,
                   motion_lanes_per_state=wshape[0], motion_obstacle_parts=wshape[1], motion_waves_per_cu=wpc,
                   validity_block=vblock, motion_block=mblock)
    try:
        env, oenv, _ = env_pair(gpu, oracle, 1002, 16, 16)
        lvs = lvs_config()
        for seed in (10, 11):
            a, b = motion_case(gpu, seed, 5000, mu=-0.5, sigma=1.0, lo=0.05, hi=5.0)
            ref_v, ref_c = oracle.check_motions(oenv, a, b, lvs, threads=8)
            A, B = to_dev(a.T.copy()), to_dev(b.T.copy())
            for _ in range(2):
                v, c = gpu.check_motions(env, A, B, lvs, exact_count=True)
                assert np.array_equal(v.cpu().numpy(), ref_v) and np.array_equal(c.cpu().numpy(), ref_c), seed
                assert np.array_equal(gpu.check_motions(env, A, B, lvs).cpu().numpy(), ref_v), seed
    finally:
        gpu.set_tuning(**DEFAULT, motion_waves_per_cu=8, validity_block=256, motion_block=256)


def test_autotune_for_streams_keeps_results(gpu, oracle):
    """The stream-pipeline autotuner tries every shape of its grid on batches
    issued round-robin on two streams, installs one of them, and the batches
    on either stream still equal the oracle; an empty stream list, or an
    environment with its own shapes, is rejected and leaves the library's
    tuning as it was."""
    import torch

    from ompl_amd._lib import OmplAmdError

    env, oenv, _ = env_pair(gpu, oracle, 1002, 24, 24)
    a, b = motion_case(gpu, 12, 2048, mu=-0.5, sigma=1.0, lo=0.05, hi=5.0)
    lvs = lvs_config()
    ref_v, _ = oracle.check_motions(oenv, a, b, lvs)
    A, B = to_dev(a.T.copy()), to_dev(b.T.copy())
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]  # default priority
    try:
        gpu.set_tuning(**DEFAULT, validity_block=512)
        t, us = gpu.autotune_motions_streams(env, A, B, lvs, streams, steps=6)
        assert us > 0 and t == gpu.get_tuning() and t["motion_strategy"] == 3 and t["validity_shared_fk"] == 0
        assert (t["validity_lanes_per_state"], t["validity_obstacle_parts"]) in \
            {(7, 1), (7, 2), (7, 4), (7, 9), (8, 4), (8, 8)}
        assert (t["motion_lanes_per_state"], t["motion_obstacle_parts"]) in {(7, 2), (7, 4), (7, 9), (8, 8)}
        assert t["motion_waves_per_cu"] in {2, 4, 8, 16}
        assert t["validity_block"] == 512  # fields outside the grid keep the tuning's values
        for s in streams:
            v = gpu.check_motions(env, A, B, lvs, stream=s.cuda_stream)
            torch.cuda.synchronize()
            assert np.array_equal(v.cpu().numpy(), ref_v)
        before = gpu.set_tuning(**DEFAULT, validity_block=256)
        with pytest.raises(OmplAmdError):
            gpu.autotune_motions_streams(env, A, B, lvs, [])
        env.set_tuning(motion_waves_per_cu=16)  # the environment's own shapes: refused
        with pytest.raises(OmplAmdError):
            gpu.autotune_motions_streams(env, A, B, lvs, streams)
        env.set_tuning(None)
        assert gpu.get_tuning() == before
    finally:
        gpu.set_tuning(**DEFAULT, motion_waves_per_cu=8, validity_block=256)


def test_env_tuning_overrides_library_tuning(gpu, oracle):
    """An environment's own launch shapes take the place of the library's for
    its batches only, and clearing them restores the library's."""
    from ompl_amd._lib import OmplAmdError

    set_strategy(gpu, 3, (8, 2), (8, 1))
    env, oenv, _ = env_pair(gpu, oracle, 1002, 6, 4)
    other, _, _ = env_pair(gpu, oracle, 1003, 6, 4)
    a, b = motion_case(gpu, 9, 3000, mu=-0.5, sigma=1.0, lo=0.05, hi=5.0)
    lvs = lvs_config()
    ref_v, ref_c = oracle.check_motions(oenv, a, b, lvs)
    A, B = to_dev(a.T.copy()), to_dev(b.T.copy())
    assert env.tuning()["own"] is False
    t = env.set_tuning(motion_strategy=2, validity_lanes_per_state=8, validity_obstacle_parts=1,
                       motion_lanes_per_state=8, motion_obstacle_parts=1)
    assert t["motion_strategy"] == 2 and env.tuning()["own"] is True
    assert gpu.get_tuning()["motion_strategy"] == 3 and other.tuning()["own"] is False

    def launches(e):
        gpu.profile.reset()
        gpu.profile.enable(True)
        try:
            v, c = gpu.check_motions(e, A, B, lvs, exact_count=True)
            return v, c, gpu.profile.query("motion_fused")[0], gpu.profile.query("motion_chunk")[0]
        finally:
            gpu.profile.enable(False)
            gpu.profile.reset()

    v, c, fused, chunk = launches(env)
    assert (fused, chunk) == (1, 0)
    assert np.array_equal(v.cpu().numpy(), ref_v) and np.array_equal(c.cpu().numpy(), ref_c)
    _, _, fused, chunk = launches(other)
    assert (fused, chunk) == (0, 1)
    with pytest.raises(OmplAmdError):
        env.set_tuning(motion_waves_per_cu=99)
    assert env.tuning()["motion_strategy"] == 2  # a rejected tuning leaves the shapes in force
    env.set_tuning(None)
    assert env.tuning()["own"] is False
    v, c, fused, chunk = launches(env)
    assert (fused, chunk) == (0, 1)
    assert np.array_equal(v.cpu().numpy(), ref_v) and np.array_equal(c.cpu().numpy(), ref_c)
    gpu.set_tuning(**DEFAULT)


def test_check_motions_edge_cases(gpu, oracle):
    env, oenv, _ = env_pair(gpu, oracle, 1003, 8, 8)
    rng = np.random.default_rng(4)
    a = rng.uniform(-PI, PI, (40, 7))
    b = a.copy()
    b[10:20] += 0.004          # nd = 1
    b[20:30] += rng.normal(size=(10, 7)) * 2.0   # long edges (nd in the thousands)
    b[30:40] = np.clip(b[30:40] + 0.9, -PI, PI)
    lvs = lvs_config()
    for factor in (1, 2):
        ref_v, ref_c = oracle.check_motions(oenv, a, b, lvs, factor)
        for tp, ops in ((3, 1), (1, 1), (0, 1), (2, 1), (4, 8), (4, 2)):
            gpu.set_tuning(**{**DEFAULT, "motion_strategy": tp, "motion_lanes_per_state": 8,
                              "motion_obstacle_parts": ops, "validity_obstacle_parts": ops})
            v, c = gpu.check_motions(env, to_dev(a.T.copy()), to_dev(b.T.copy()), lvs, factor, exact_count=True)
            assert np.array_equal(v.cpu().numpy(), ref_v)
            assert np.array_equal(c.cpu().numpy(), ref_c)
    gpu.set_tuning(**DEFAULT)


def test_check_motion_plain_single_call(gpu, oracle):
    env, oenv, _ = env_pair(gpu, oracle, 1002, 6, 4)
    a, b = motion_case(gpu, 12, 300, mu=-0.5)
    lvs = lvs_config()
    for i in range(a.shape[0]):
        assert gpu.check_motion(env, a[i], b[i], lvs) == oracle.check_motion(oenv, a[i], b[i], lvs)[0]


@pytest.mark.parametrize("lps,ops", [(1, 1), (8, 1), (8, 8), (7, 9), (7, 2)])
def test_check_motions_last_valid(gpu, oracle, lps, ops):
    gpu.set_tuning(motion_lanes_per_state=lps, motion_obstacle_parts=ops)
    env, oenv, _ = env_pair(gpu, oracle, 1002, 6, 4)
    a, b = motion_case(gpu, 13, 2000, mu=-0.5)
    lvs = lvs_config()
    v, t = gpu.check_motions_last_valid(env, to_dev(a.T.copy()), to_dev(b.T.copy()), lvs)
    v, t = v.cpu().numpy(), t.cpu().numpy()
    for i in range(a.shape[0]):
        ov, ot, _ = oracle.check_motion_last_valid(oenv, a[i], b[i], lvs)
        assert v[i] == ov
        if not ov:
            assert t[i] == ot, (i, t[i], ot)
        else:
            assert math.isnan(t[i])
    for i in range(0, 200, 7):
        ok, tt = gpu.check_motion_last_valid(env, a[i], b[i], lvs)
        ov, ot, _ = oracle.check_motion_last_valid(oenv, a[i], b[i], lvs)
        assert ok == ov and (ov or tt == ot)
    gpu.set_tuning(**DEFAULT)


def test_empty_batches(gpu):
    import torch

    env = gpu.Env(np.zeros((1, 4), np.float32), None)
    z = torch.empty((7, 0), dtype=torch.float64, device="cuda")
    assert gpu.check_motions(env, z, z, 0.01).numel() == 0
    assert gpu.validate_states(env, z.float()).numel() == 0


@pytest.mark.parametrize("counts", [True, False])
def test_motion_graph_group_replays(gpu, oracle, counts):
    """Three independent batches of different sizes captured in one graph on
    forked streams: every replay gives each batch's eager result, with eager
    batches in between."""
    import torch

    set_strategy(gpu, 3, (7, 4), (7, 9))
    env, oenv, _ = env_pair(gpu, oracle, 1001, 128, 128)
    lvs = lvs_config()
    cases = [motion_case(gpu, 11 + i, n) for i, n in enumerate((4096, 1000, 2500))]
    refs = [oracle.check_motions(oenv, a, b, lvs) for a, b in cases]
    batches = [(to_dev(a.T.copy()), to_dev(b.T.copy())) for a, b in cases]
    checked = [torch.empty(a.shape[0], dtype=torch.int32, device="cuda") for a, _ in cases] if counts else None
    g = gpu.MotionGraph.group(env, batches, lvs, checked=checked)
    a2, b2 = motion_case(gpu, 5, 2048)
    A2, B2 = to_dev(a2.T.copy()), to_dev(b2.T.copy())
    ref2, _ = oracle.check_motions(oenv, a2, b2, lvs)
    for rep in range(5):
        for o in g.out:
            o.fill_(7)
        if counts:
            for c in checked:
                c.fill_(-1)
        g.launch()
        torch.cuda.synchronize()
        for i, (rv, rc) in enumerate(refs):
            assert np.array_equal(g.out[i].cpu().numpy(), rv), (rep, i)
            if counts:
                assert np.array_equal(checked[i].cpu().numpy(), rc), (rep, i)
        v2 = gpu.check_motions(env, A2, B2, lvs)
        torch.cuda.synchronize()
        assert np.array_equal(v2.cpu().numpy(), ref2), rep
    g.close()
    gpu.set_tuning(**DEFAULT)


@pytest.mark.parametrize("strategy,eshape,wshape", [(4, (8, 8), (8, 8)), (3, (8, 2), (8, 1)), (3, (7, 4), (7, 9)),
                                                    (1, (8, 1), (1, 1)), (0, (8, 1), (8, 8)), (2, (8, 2), (8, 8))])
@pytest.mark.parametrize("counts", [True, False])
def test_motion_graph_replays_match_eager(gpu, oracle, strategy, eshape, wshape, counts):
    """A captured batch replayed several times, with eager batches and host
    synchronisations in between, gives the eager result on every replay --
    with checked counts (a results kernel in the graph) and without (the
    replays alternate between the batch's two recordings, one per survivor
    counter set)."""
    import torch

    set_strategy(gpu, strategy, eshape, wshape)
    env, oenv, _ = env_pair(gpu, oracle, 1001, 128, 128)
    a, b = motion_case(gpu, 4, 4096)
    lvs = lvs_config()
    ref_v, ref_c = oracle.check_motions(oenv, a, b, lvs)
    A, B = to_dev(a.T.copy()), to_dev(b.T.copy())
    checked = torch.empty(4096, dtype=torch.int32, device="cuda") if counts else None
    g = gpu.MotionGraph(env, A, B, lvs, checked=checked)
    a2, b2 = motion_case(gpu, 5, 2048)
    A2, B2 = to_dev(a2.T.copy()), to_dev(b2.T.copy())
    ref2, _ = oracle.check_motions(oenv, a2, b2, lvs)
    for rep in range(5):
        g.out.fill_(7)
        if counts:
            checked.fill_(-1)
        g.launch()
        torch.cuda.synchronize()
        assert np.array_equal(g.out.cpu().numpy(), ref_v), rep
        if counts:
            assert np.array_equal(checked.cpu().numpy(), ref_c), rep
        v2 = gpu.check_motions(env, A2, B2, lvs)  # eager work between replays
        torch.cuda.synchronize()
        assert np.array_equal(v2.cpu().numpy(), ref2), rep
    g.close()
    gpu.set_tuning(**DEFAULT)


@pytest.mark.parametrize("strategy,eshape,wshape", [(3, (8, 8), (8, 4)), (4, (8, 2), (8, 8)), (1, (8, 4), (1, 1)),
                                                    (3, (8, 4), (1, 1))])
def test_check_motions_shared_s2_pass(gpu, oracle, strategy, eshape, wshape):
    """The s2 pass with shared FK (validity_shared_fk) gives the same booleans and counts."""
    set_strategy(gpu, strategy, eshape, wshape)
    gpu.set_tuning(validity_shared_fk=1)
    env, oenv, _ = env_pair(gpu, oracle, 1001, 128, 128)
    a, b = motion_case(gpu, 4, 16384)
    lvs = lvs_config()
    ref_v, ref_c = oracle.check_motions(oenv, a, b, lvs)
    v, c = gpu.check_motions(env, to_dev(a.T.copy()), to_dev(b.T.copy()), lvs, exact_count=True)
    assert np.array_equal(v.cpu().numpy(), ref_v)
    assert np.array_equal(c.cpu().numpy(), ref_c)
    gpu.set_tuning(**DEFAULT, validity_shared_fk=0)


def test_capabilities_match_enforced_limits(gpu):
    from ompl_amd._lib import OmplAmdError

    caps = gpu.capabilities()
    assert caps["arm_dof"] == 7 and caps["arm_spheres"] == 28 and caps["compute_units"] >= 1
    assert caps["device_name"].startswith("gfx950")
    w = caps["max_env_words"]
    gpu.Env(np.zeros((w - 2, 4), np.float32), np.zeros((1, 6), np.float32)).close()  # exactly at the limit
    with pytest.raises(OmplAmdError):
        gpu.Env(np.zeros((w - 1, 4), np.float32), np.zeros((1, 6), np.float32))


def test_largest_environment_every_batched_path(gpu, oracle):
    """An environment at the obstacle limit (64 KiB of obstacles in LDS, beside
    the kernels' own LDS): validity, motions, lastValid and state lists match
    the oracle.  Most obstacles are far from the arm, a few near ones decide."""
    import torch

    w = gpu.capabilities()["max_env_words"]
    rng = np.random.default_rng(21)
    s = np.zeros((w - 2, 4), np.float32)
    s[:, :3] = rng.uniform(20.0, 30.0, (w - 2, 3))
    s[:, 3] = 0.1
    near, _ = gpu.workload.obstacles(1004, 12, 0)
    s[:12] = near
    b = np.array([[25.0, 25.0, 25.0, 0.5, 0.5, 0.5]], np.float32)
    env, oenv = gpu.Env(s, b), oracle.OracleEnv(s, b)
    q = gpu.workload.configs(31, 256)
    ref = np.array([oracle.state_valid(oenv, r.astype(np.float32)) for r in q])
    assert 0 < ref.sum() < len(ref)
    got = gpu.validate_states(env, to_dev(q.T.astype(np.float32).copy())).cpu().numpy()
    assert np.array_equal(got.astype(bool), ref)
    a, bb = motion_case(gpu, 32, 128, mu=-1.5)
    lvs = lvs_config()
    ref_v, ref_c = oracle.check_motions(oenv, a, bb, lvs)
    A, B = to_dev(a.T.copy()), to_dev(bb.T.copy())
    for strategy, e, wsh in ((3, (8, 2), (8, 1)), (1, (8, 1), (1, 1)), (3, (7, 4), (7, 9))):
        set_strategy(gpu, strategy, e, wsh)
        v, c = gpu.check_motions(env, A, B, lvs, exact_count=True)
        assert np.array_equal(v.cpu().numpy(), ref_v) and np.array_equal(c.cpu().numpy(), ref_c), strategy
        lv, lt = gpu.check_motions_last_valid(env, A, B, lvs)
        lv, lt = lv.cpu().numpy(), lt.cpu().numpy()
        for i in range(0, 128, 9):
            ov, ot, _ = oracle.check_motion_last_valid(oenv, a[i], bb[i], lvs)
            assert lv[i] == ov and (ov or lt[i] == ot), (strategy, i)
        states = np.concatenate([q[:40], q[40:80]])
        off = np.array([0, 40, 80], np.int64)
        sv, sf = gpu.check_state_lists(env, to_dev(states.T.copy()), torch.from_numpy(off).cuda())
        for i in range(2):
            ok, f = oracle.check_states_first_invalid(oenv, states[off[i]:off[i + 1]])
            assert bool(sv[i].item()) == ok and int(sf[i].item()) == (-1 if ok else f)
    gpu.set_tuning(**DEFAULT)
    env.close()


def test_counter_sets_across_mixed_calls(gpu, oracle):
    """Batches with and without checked counts, captured-graph replays, a batch
    with no survivors and batches of other sizes, in one sequence on one stream
    (the walk's counter sets alternate only for batches without a results
    kernel): every result equals the oracle's."""
    import torch

    gpu.set_tuning(motion_strategy=3, validity_lanes_per_state=7, validity_obstacle_parts=4,
                   motion_lanes_per_state=7, motion_obstacle_parts=9)
    env, oenv, _ = env_pair(gpu, oracle, 1001, 128, 128)
    lvs = lvs_config()
    cases = []
    for seed, n in ((4, 4096), (5, 1000), (6, 77)):
        a, b = motion_case(gpu, seed, n)
        cases.append((to_dev(a.T.copy()), to_dev(b.T.copy()), *oracle.check_motions(oenv, a, b, lvs)))
    # a batch whose every s2 collides: no survivors (s2 at the all-zero configuration inside an obstacle is
    # not guaranteed, so take edges ending in states the oracle marks invalid)
    a, b = motion_case(gpu, 7, 3000)
    s2_ok = oracle.validate_states(oenv, np.ascontiguousarray(b.T.astype(np.float32)))
    bad = np.nonzero(s2_ok == 0)[0][:500]
    a0, b0 = a[bad], b[bad]
    cases.append((to_dev(a0.T.copy()), to_dev(b0.T.copy()), *oracle.check_motions(oenv, a0, b0, lvs)))
    A0, B0, ref_v0, ref_c0 = cases[0]
    g = gpu.MotionGraph(env, A0, B0, lvs)
    for rep in range(3):
        for i, (A, B, ref_v, ref_c) in enumerate(cases):
            print(f"mixed calls rep {rep} case {i}", flush=True)
            v = gpu.check_motions(env, A, B, lvs)
            assert np.array_equal(v.cpu().numpy(), ref_v), (rep, i)
            if i % 2 == rep % 2:
                v, c = gpu.check_motions(env, A, B, lvs, exact_count=True)
                assert np.array_equal(v.cpu().numpy(), ref_v) and np.array_equal(c.cpu().numpy(), ref_c), (rep, i)
            g.launch()
            torch.cuda.synchronize()
            assert np.array_equal(g.out.cpu().numpy(), ref_v0), (rep, i)
    g.close()
    gpu.set_tuning(**DEFAULT)


def test_two_streams_concurrently(gpu, oracle):
    """Batches issued round-robin on two streams (each with its own survivor
    workspace) give the oracle's results."""
    import torch

    gpu.set_tuning(motion_strategy=3, validity_lanes_per_state=7, validity_obstacle_parts=4,
                   motion_lanes_per_state=7, motion_obstacle_parts=9)
    env, oenv, _ = env_pair(gpu, oracle, 1001, 128, 128)
    lvs = lvs_config()
    a1, b1 = motion_case(gpu, 4, 8192)
    a2, b2 = motion_case(gpu, 8, 5000)
    r1, _ = oracle.check_motions(oenv, a1, b1, lvs)
    r2, _ = oracle.check_motions(oenv, a2, b2, lvs)
    A1, B1, A2, B2 = to_dev(a1.T.copy()), to_dev(b1.T.copy()), to_dev(a2.T.copy()), to_dev(b2.T.copy())
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    o1 = [torch.empty(8192, dtype=torch.uint8, device="cuda") for _ in range(4)]
    o2 = [torch.empty(5000, dtype=torch.uint8, device="cuda") for _ in range(4)]
    for i in range(4):
        with torch.cuda.stream(s1):
            gpu.check_motions(env, A1, B1, lvs, out=o1[i])
        with torch.cuda.stream(s2):
            gpu.check_motions(env, A2, B2, lvs, out=o2[i])
    torch.cuda.synchronize()
    for i in range(4):
        assert np.array_equal(o1[i].cpu().numpy(), r1) and np.array_equal(o2[i].cpu().numpy(), r2), i
    gpu.set_tuning(**DEFAULT)


def test_two_threads_share_one_stream(gpu, oracle):
    """Boolean-only batches issued by two host threads on one stream share that
    stream's survivor workspace: each batch takes its counter set and enqueues
    its launches under the workspace's lock, so every batch gives the oracle's
    results (without the lock two batches could pick the same set)."""
    import threading

    import torch

    gpu.set_tuning(motion_strategy=3, validity_lanes_per_state=7, validity_obstacle_parts=4,
                   motion_lanes_per_state=7, motion_obstacle_parts=9)
    env, oenv, _ = env_pair(gpu, oracle, 1001, 128, 128)
    lvs = lvs_config()
    cases = [motion_case(gpu, 4, 8192), motion_case(gpu, 8, 3000)]
    refs = [oracle.check_motions(oenv, a, b, lvs)[0] for a, b in cases]
    dev = [(to_dev(a.T.copy()), to_dev(b.T.copy())) for a, b in cases]
    shared = torch.cuda.Stream()
    outs = {0: [], 1: []}
    errors = []

    def work(t):
        try:
            with torch.cuda.stream(shared):
                for i in range(12):
                    A, B = dev[(i + t) % 2]
                    outs[t].append(((i + t) % 2, gpu.check_motions(env, A, B, lvs)))
        except Exception as e:  # reported below
            errors.append(e)

    th = [threading.Thread(target=work, args=(t,)) for t in (0, 1)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    torch.cuda.synchronize()
    assert not errors, errors
    for t in (0, 1):
        for i, (c, o) in enumerate(outs[t]):
            assert np.array_equal(o.cpu().numpy(), refs[c]), (t, i)
    gpu.set_tuning(**DEFAULT)
