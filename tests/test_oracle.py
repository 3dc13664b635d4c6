"""Pins the CPU oracle against the reference's golden data, known answers and
closed forms (CPU only).  Each test is chosen so that a plausible slip in the
oracle (a dropped term, a wrong sign, a transposed operand, an off-by-one in
the bisection queue) fails it."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"
PI = math.pi


# ---------------------------------------------------------------- transcendentals
def test_sincos_accuracy(oracle):
    xs = np.concatenate([np.linspace(-PI, PI, 20001), np.linspace(-20, 20, 4001)]).astype(np.float32)
    err_s = err_c = 0.0
    for x in xs:
        s, c = oracle.sincosf(float(x))
        err_s = max(err_s, abs(s - math.sin(float(x))))
        err_c = max(err_c, abs(c - math.cos(float(x))))
    assert err_s < 3e-7 and err_c < 3e-7, (err_s, err_c)
    assert oracle.sincosf(0.0) == (0.0, 1.0)


def test_sincos_quadrants(oracle):
    # sin(x + pi/2) = cos(x) etc. pins the quadrant table
    for x in np.linspace(-1.2, 1.2, 97):
        s0, c0 = math.sin(x), math.cos(x)
        for k, (es, ec) in enumerate([(s0, c0), (c0, -s0), (-s0, -c0), (-c0, s0)]):
            s, c = oracle.sincosf(float(np.float32(x + k * PI / 2)))
            assert abs(s - es) < 1e-6 and abs(c - ec) < 1e-6


def test_acos_accuracy(oracle):
    xs = np.linspace(0, 1, 50001).astype(np.float32)
    err = max(abs(oracle.acosf(float(x)) - math.acos(float(x))) for x in xs)
    assert err < 4e-7, err
    assert oracle.acosf(1.0) == 0.0


# ---------------------------------------------------------------- arm FK
def rz(t):
    c, s = math.cos(t), math.sin(t)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def ry(t):
    c, s = math.cos(t), math.sin(t)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def fk_fp64(q):
    """Independent FP64 restatement with rotation matrices."""
    R = np.eye(3)
    p = np.zeros(3)
    out = []
    for i, qi in enumerate(q):
        R = R @ (rz(qi) if i % 2 == 0 else ry(qi))
        z = R[:, 2]
        for f in (0.125, 0.375, 0.625, 0.875):
            out.append(p + 0.3 * f * z)
        p = p + 0.3 * z
    return np.array(out)


def test_fk_zero_config(oracle):
    c = oracle.arm_spheres(np.zeros(7))
    assert np.all(c[:, 0] == 0) and np.all(c[:, 1] == 0)
    expect = np.array([0.3 * i + 0.3 * f for i in range(7) for f in (0.125, 0.375, 0.625, 0.875)])
    assert np.max(np.abs(c[:, 2] - expect)) < 1e-6


def test_fk_single_joint_closed_form(oracle):
    # joint 1 (y axis) alone tilts links 1..6 by theta in the x-z plane
    for th in (-2.0, -0.5, 0.3, 1.0, 3.0):
        q = np.zeros(7)
        q[1] = th
        c = oracle.arm_spheres(q)
        for li in range(1, 7):
            for j, f in enumerate((0.125, 0.375, 0.625, 0.875)):
                dist = 0.3 * (li - 1) + 0.3 * f
                exp = np.array([math.sin(th) * dist, 0.0, 0.3 + math.cos(th) * dist])
                assert np.max(np.abs(c[li * 4 + j] - exp)) < 2e-6


def test_fk_matches_fp64_rotation_matrices(oracle):
    rng = np.random.default_rng(1)
    for _ in range(300):
        q = rng.uniform(-PI, PI, 7)
        c = oracle.arm_spheres(q.astype(np.float32))
        ref = fk_fp64(q.astype(np.float32).astype(np.float64))
        assert np.max(np.abs(c - ref)) < 5e-6


# ---------------------------------------------------------------- collision
def classify_fp64(q, spheres, boxes):
    """min margin over all pairs in FP64 (negative = collision)."""
    cs = fk_fp64(np.asarray(q, np.float32).astype(np.float64))
    m = np.inf
    if len(spheres):
        d = np.linalg.norm(cs[:, None, :] - spheres[None, :, :3].astype(np.float64), axis=2)
        m = min(m, np.min(d - (spheres[None, :, 3].astype(np.float64) + 0.05)))
    if len(boxes):
        c = boxes[:, :3].astype(np.float64)
        h = boxes[:, 3:].astype(np.float64)
        p = np.clip(cs[:, None, :], c - h, c + h)
        d = np.linalg.norm(cs[:, None, :] - p, axis=2)
        m = min(m, np.min(d - 0.05))
    return m


def random_env(rng, ns, nb):
    def centres(n):
        out = []
        while len(out) < n:
            c = rng.uniform(-1, 1, 3)
            if np.linalg.norm(c) >= 0.2:
                out.append(c)
        return np.array(out, np.float64).reshape(-1, 3)

    s = np.zeros((ns, 4), np.float32)
    s[:, :3] = centres(ns)
    s[:, 3] = rng.uniform(0.05, 0.15, ns)
    b = np.zeros((nb, 6), np.float32)
    b[:, :3] = centres(nb)
    b[:, 3:] = rng.uniform(0.05, 0.2, (nb, 3))
    return s, b


@pytest.mark.parametrize("ns,nb", [(16, 0), (0, 16), (24, 24)])
def test_collision_matches_fp64_brute_force(oracle, ns, nb):
    rng = np.random.default_rng(ns * 100 + nb)
    s, b = random_env(rng, ns, nb)
    env = oracle.OracleEnv(s, b)
    agree = amb = 0
    for _ in range(400):
        q = rng.uniform(-PI, PI, 7).astype(np.float32)
        m = classify_fp64(q, s, b)
        v = oracle.state_valid(env, q)
        if abs(m) < 1e-5:
            amb += 1
            continue
        assert v == (m > 0), (q, m)
        agree += 1
    assert agree > 350


def test_collision_known_answers(oracle):
    # zero config: sphere 0 of link 0 at (0, 0, 0.0375)
    far = oracle.OracleEnv(np.array([[5.0, 5.0, 5.0, 0.1]]), np.zeros((0, 6)))
    assert oracle.state_valid(far, np.zeros(7))
    on_top = oracle.OracleEnv(np.array([[0.0, 0.0, 0.0375, 0.01]]), np.zeros((0, 6)))
    assert not oracle.state_valid(on_top, np.zeros(7))
    box_around_tip = oracle.OracleEnv(np.zeros((0, 4)), np.array([[0.0, 0.0, 2.0, 0.1, 0.1, 0.1]]))
    assert not oracle.state_valid(box_around_tip, np.zeros(7))  # tip sphere centre at z = 1.8+0.2625
    box_beside = oracle.OracleEnv(np.zeros((0, 4)), np.array([[0.5, 0.0, 1.0, 0.1, 0.1, 0.1]]))
    assert oracle.state_valid(box_beside, np.zeros(7))


# ---------------------------------------------------------------- DMV pieces
def test_longest_valid_segment_and_nd(oracle):
    lo, hi = [-PI] * 7, [PI] * 7
    ext = oracle.max_extent(lo, hi)
    assert abs(ext - 2 * PI * math.sqrt(7)) < 1e-12
    lvs = oracle.longest_valid_segment(ext, 0.01 / ext)
    assert abs(lvs - 0.01) < 1e-15
    rng = np.random.default_rng(3)
    for _ in range(200):
        a = rng.uniform(-PI, PI, 7)
        b = a + rng.normal(size=7) * rng.uniform(0, 1)
        d = math.sqrt(sum((x - y) ** 2 for x, y in zip(a, b)))
        assert oracle.valid_segment_count(a, b, lvs, 1) == math.ceil(d / lvs)
        assert oracle.valid_segment_count(a, b, lvs, 3) == 3 * math.ceil(d / lvs)
    assert oracle.valid_segment_count(np.zeros(7), np.zeros(7), lvs, 1) == 0


def test_bfs_order_known_answers(oracle):
    assert list(oracle.bfs_order(2)) == [1]
    assert list(oracle.bfs_order(3)) == [1, 2]
    assert list(oracle.bfs_order(5)) == [2, 1, 3, 4]
    assert list(oracle.bfs_order(8)) == [4, 2, 6, 1, 3, 5, 7]
    assert list(oracle.bfs_order(1)) == []


def test_bfs_order_is_bisection(oracle):
    for nd in list(range(2, 300)) + [511, 512, 513, 1000, 4097]:
        order = oracle.bfs_order(nd)
        assert sorted(order) == list(range(1, nd))
        # after the first 2^l - 1 checks the largest unchecked gap has shrunk
        # geometrically: bisection, not a scan
        for l in range(1, 12):
            cnt = (1 << l) - 1
            if cnt >= nd - 1:
                break
            pts = sorted([0, nd] + list(order[:cnt]))
            gap = max(b - a for a, b in zip(pts, pts[1:]))
            assert gap <= math.ceil(nd / (1 << l)) + 1, (nd, l, gap)


def bisection_index(p, nd):
    """Closed-form position of the p-th bisection check (the algorithm of the
    GPU kernels), checked here against the oracle's FIFO simulation."""
    m = nd - 1
    L = m.bit_length() - 1
    a, b = 1, m
    full = (1 << L) - 1
    if p < full:
        l = (p + 1).bit_length() - 1
        i = p + 1 - (1 << l)
        for t in range(l - 1, -1, -1):
            mid = (a + b) >> 1
            if (i >> t) & 1:
                a = mid + 1
            else:
                b = mid - 1
        return (a + b) >> 1
    j = p - full
    for h in range(L, 0, -1):
        mid = (a + b) >> 1
        s = b - a + 1
        left = ((s - 1) >> 1) - ((1 << (h - 1)) - 1)
        if j < left:
            b = mid - 1
        else:
            j -= left
            a = mid + 1
    return (a + b) >> 1


def test_closed_form_bisection_matches_queue(oracle):
    for nd in list(range(2, 1200)) + [2047, 2048, 2049, 4095, 4096, 10000]:
        order = oracle.bfs_order(nd)
        assert [bisection_index(p, nd) for p in range(nd - 1)] == list(order), nd


def test_check_motion_closed_form(oracle):
    rng = np.random.default_rng(11)
    s, b = random_env(rng, 12, 12)
    env = oracle.OracleEnv(s, b)
    lvs = 0.05
    n_valid = n_bad = 0
    for _ in range(150):
        a = rng.uniform(-PI, PI, 7)
        e = np.clip(a + rng.normal(size=7) * 0.3, -PI, PI)
        nd = oracle.valid_segment_count(a, e, lvs, 1)
        inter = [oracle.state_valid(env, (a + (e - a) * (j / nd)).astype(np.float32)) for j in range(1, nd)]
        end_ok = oracle.state_valid(env, e.astype(np.float32))
        expect = end_ok and all(inter)
        v, checked = oracle.check_motion(env, a, e, lvs, 1)
        assert v == expect
        # checked count: s2 first, then bisection order up to the first invalid
        if not end_ok:
            assert checked == 1
        else:
            order = oracle.bfs_order(nd)
            bad = [k for k, j in enumerate(order) if not inter[j - 1]]
            assert checked == (1 + len(order) if not bad else 2 + bad[0])
        # lastValid overload: sequential scan, closed form (j-1)/nd
        v2, t, chk2 = oracle.check_motion_last_valid(env, a, e, lvs, 1)
        assert v2 == expect
        if not expect:
            first = next((j for j in range(1, nd) if not inter[j - 1]), None)
            exp_t = (first - 1) / nd if first is not None else (nd - 1) / nd
            assert t == exp_t
        n_valid += v
        n_bad += not v
    assert n_valid > 5 and n_bad > 5


# ---------------------------------------------------------------- RNG
def test_mt19937_known_answers(oracle):
    g = json.loads((GOLDEN / "stdlib_random.json").read_text())
    m = oracle.Mt19937(5489)
    assert [m.next() for _ in range(32)] == g["mt_default_first"]
    m = oracle.Mt19937(5489)
    v = 0
    for _ in range(10000):
        v = m.next()
    assert v == 4123659995 == g["mt_default_10000th"]  # C++ [rand.predef]


def test_uniform_real_matches_libstdcxx(oracle):
    g = json.loads((GOLDEN / "stdlib_random.json").read_text())
    for seed, vals in g["uniform01"].items():
        m = oracle.Mt19937(int(seed))
        assert [m.uniform_real(0.0, 1.0) for _ in vals] == vals
    m = oracle.Mt19937(7)
    got = m.sample_uniform([-PI] * 7, [PI] * 7, 8).ravel().tolist()
    assert got == g["sample_uniform_seed7_pi"]


def test_quaternion_unit_norm(oracle):
    m = oracle.Mt19937(3)
    for _ in range(500):
        q = m.quaternion()
        assert abs(np.linalg.norm(q) - 1.0) < 1e-12


# ---------------------------------------------------------------- NN
def test_se3_distance_closed_forms(oracle):
    ident = np.array([0, 0, 0, 0, 0, 0, 1], np.float32)
    t = np.array([3, 4, 0, 0, 0, 0, 1], np.float32)
    assert oracle.se3_dist(t, ident) == 5.0
    for th in (0.1, 0.7, 1.5, 2.5):
        q = np.array([0, 0, 0, 0, 0, math.sin(th / 2), math.cos(th / 2)], np.float32)
        assert abs(oracle.se3_dist(q, ident) - th / 2) < 2e-6  # arc of the half angle
    rng = np.random.default_rng(5)
    for _ in range(200):
        a = np.concatenate([rng.uniform(-1, 1, 3), rng.normal(size=4)])
        b = np.concatenate([rng.uniform(-1, 1, 3), rng.normal(size=4)])
        a[3:] /= np.linalg.norm(a[3:])
        b[3:] /= np.linalg.norm(b[3:])
        a, b = a.astype(np.float32), b.astype(np.float32)
        assert oracle.se3_dist(a, b) == oracle.se3_dist(b, a)
        # acos(|q.q|) near 1 is ill-conditioned: an FP32-rounded unit quaternion
        # has |q|^2 = 1 - O(1e-7), so d(a, a) = acos(1 - d) ~ sqrt(2d) <= ~5e-4,
        # in FP32 and in the reference's FP64 alike (its 1e-9 cut-off,
        # base/spaces/src/SO3StateSpace.cpp:258, is below FP32 rounding)
        self_ref = math.acos(min(1.0, abs(float(np.dot(a[3:].astype(np.float64), a[3:])))))
        assert oracle.se3_dist(a, a) <= 6e-4 and abs(oracle.se3_dist(a, a) - self_ref) < 4e-4
        ref = np.linalg.norm(a[:3].astype(np.float64) - b[:3]) + math.acos(min(1.0, abs(np.dot(a[3:].astype(
            np.float64), b[3:]))))
        assert abs(oracle.se3_dist(a, b) - ref) < 1e-5


def _near_tie(d_sorted_ref, k, tol):
    return k < len(d_sorted_ref) and abs(d_sorted_ref[k] - d_sorted_ref[k - 1]) < tol


@pytest.mark.parametrize("name,metric", [("nn_linear_l2.npz", 0), ("nn_linear_se3.npz", 1)])
def test_oracle_knn_matches_reference_linear(oracle, name, metric):
    """Golden data of the reference's own NearestNeighborsLinear (FP64)."""
    f = np.load(GOLDEN / name)
    data, queries, k = f["data"], f["queries"], int(f["k"])
    idx, dist = oracle.knn(metric, data, queries, k)
    n_tie = 0
    for j in range(queries.shape[0]):
        ref_i, ref_d = f["knn_idx"][j], f["knn_dist"][j]
        assert np.max(np.abs(dist[j] - ref_d)) < 2e-5
        # full FP64 distance list to detect a near tie at the k-th boundary
        if metric == 0:
            all_d = np.sqrt(((data.astype(np.float64) - queries[j].astype(np.float64)) ** 2).sum(1))
        else:
            all_d = np.array([oracle.se3_dist(e, queries[j]) for e in data], np.float64)
        srt = np.sort(all_d)
        if _near_tie(srt, k, 1e-5):
            n_tie += 1
            continue
        assert set(idx[j]) == set(ref_i.tolist())
    assert n_tie <= 2


@pytest.mark.parametrize("name,metric", [("nn_linear_l2.npz", 0), ("nn_linear_se3.npz", 1)])
def test_oracle_radius_and_nearest_match_reference_linear(oracle, name, metric):
    f = np.load(GOLDEN / name)
    data, queries, r = f["data"], f["queries"], float(f["radius"])
    res = oracle.radius(metric, data, queries, r)
    off = f["rad_off"]
    for j, (ii, dd) in enumerate(res):
        ref_i = f["rad_idx"][off[j]:off[j + 1]]
        ref_d = f["rad_dist"][off[j]:off[j + 1]]
        diff = set(ii.tolist()) ^ set(ref_i.tolist())
        # only boundary elements (|d - r| within FP32 error) may differ
        for e in diff:
            de = oracle.l2_dist(data[e], queries[j]) if metric == 0 else oracle.se3_dist(data[e], queries[j])
            assert abs(de - r) < 1e-5
        assert np.all(np.diff(dd) >= 0)
        if not diff:
            assert np.max(np.abs(dd - ref_d), initial=0) < 2e-5
        assert oracle.nearest(metric, data, queries[j]) == f["nearest"][j]


def test_distance_and_interpolation_invariants(oracle):
    """The reference's own state-space invariants (tests/base/StateSpaceTest.h:71-117,
    eps = 10 * DBL_EPSILON): d(s, s) = 0, d > 0 and symmetric for distinct states,
    interpolation at 0 / 1 gives the endpoints, the midpoint satisfies the triangle
    equality, and interpolate(interp(s3, s2, .5)) == interpolate(s1, s2, .75) for
    s3 = interp(s1, s2, .5)."""
    eps = 10 * np.finfo(np.float64).eps
    rng = np.random.default_rng(17)
    for _ in range(200):
        s1, s2 = rng.uniform(-np.pi, np.pi, (2, 7))
        assert abs(oracle.distance(s1, s1)) <= eps
        d12 = oracle.distance(s1, s2)
        assert d12 > 0 and abs(d12 - oracle.distance(s2, s1)) <= eps
        assert oracle.distance(s1, oracle.interpolate(s1, s2, 0.0)) <= eps
        assert oracle.distance(s2, oracle.interpolate(s1, s2, 1.0)) <= eps
        s3 = oracle.interpolate(s1, s2, 0.5)
        assert abs(oracle.distance(s1, s3) + oracle.distance(s3, s2) - d12) <= eps * max(1.0, d12)
        s3b = oracle.interpolate(s3, s2, 0.5)
        s2b = oracle.interpolate(s1, s2, 0.75)
        assert oracle.distance(s2b, s3b) <= eps
