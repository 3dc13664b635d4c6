"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never by the product.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"
REF_LINEAR = HERE / "_ref" / "ref_linear"
REFERENCE_SRC = Path("/root/reference/src")

vp = C.c_void_p


def build(force: bool = False) -> Path:
    """Builds the oracle (and the reference Linear driver when the reference tree exists)."""
    if force or not LIB.exists():
        subprocess.run(["make", "-C", str(HERE), "all"], check=True, capture_output=True)
    if REFERENCE_SRC.exists() and (force or not REF_LINEAR.exists()):
        subprocess.run(["make", "-C", str(HERE), "ref"], check=True, capture_output=True)
    return LIB


class Env(C.Structure):
    _fields_ = [("n_spheres", C.c_int), ("spheres", vp), ("n_boxes", C.c_int), ("boxes", vp)]


class MT(C.Structure):
    _fields_ = [("mt", C.c_uint32 * 624), ("idx", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        L.orc_sincosf.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.orc_acosf.argtypes = [C.c_float]
        L.orc_acosf.restype = C.c_float
        L.orc_arm_spheres.argtypes = [vp, vp]
        L.orc_state_valid.argtypes = [C.POINTER(Env), vp, C.POINTER(C.c_long)]
        L.orc_state_valid_d.argtypes = [C.POINTER(Env), vp, C.POINTER(C.c_long)]
        L.orc_validate_states.argtypes = [C.POINTER(Env), vp, C.c_long, vp]
        L.orc_max_extent.argtypes = [vp, vp, C.c_int]
        L.orc_max_extent.restype = C.c_double
        L.orc_longest_valid_segment.argtypes = [C.c_double, C.c_double]
        L.orc_longest_valid_segment.restype = C.c_double
        L.orc_distance_d.argtypes = [vp, vp, C.c_int]
        L.orc_distance_d.restype = C.c_double
        L.orc_interpolate_d.argtypes = [vp, vp, C.c_double, C.c_int, vp]
        L.orc_valid_segment_count.argtypes = [vp, vp, C.c_int, C.c_double, C.c_uint]
        L.orc_valid_segment_count.restype = C.c_uint
        L.orc_bfs_order.argtypes = [C.c_int, vp]
        L.orc_check_motion.argtypes = [C.POINTER(Env), vp, vp, C.c_double, C.c_uint, C.POINTER(C.c_int)]
        L.orc_check_motion_last_valid.argtypes = [C.POINTER(Env), vp, vp, C.c_double, C.c_uint,
                                                  C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.orc_check_states.argtypes = [C.POINTER(Env), vp, C.c_int, C.POINTER(C.c_int)]
        L.orc_check_states_first_invalid.argtypes = [C.POINTER(Env), vp, C.c_int, C.POINTER(C.c_int)]
        L.orc_sample_valid.argtypes = [C.POINTER(Env), C.POINTER(MT), vp, vp, C.c_long, C.c_int, vp, vp, vp]
        L.orc_mt_seed.argtypes = [C.POINTER(MT), C.c_uint32]
        L.orc_mt_next.argtypes = [C.POINTER(MT)]
        L.orc_mt_next.restype = C.c_uint32
        L.orc_canonical.argtypes = [C.POINTER(MT)]
        L.orc_canonical.restype = C.c_double
        L.orc_uniform_real.argtypes = [C.POINTER(MT), C.c_double, C.c_double]
        L.orc_uniform_real.restype = C.c_double
        L.orc_sample_uniform.argtypes = [C.POINTER(MT), vp, vp, C.c_int, C.c_long, vp]
        L.orc_quaternion.argtypes = [C.POINTER(MT), vp]
        L.orc_l2_dist.argtypes = [vp, vp, C.c_int]
        L.orc_l2_dist.restype = C.c_float
        L.orc_se3_dist.argtypes = [vp, vp]
        L.orc_se3_dist.restype = C.c_float
        L.orc_knn.argtypes = [C.c_int, vp, C.c_long, C.c_int, vp, C.c_int, vp, vp]
        L.orc_radius.argtypes = [C.c_int, vp, C.c_long, C.c_int, vp, C.c_float, vp, vp, C.c_long]
        L.orc_radius.restype = C.c_long
        L.orc_nearest.argtypes = [C.c_int, vp, C.c_long, C.c_int, vp]
        L.orc_nearest.restype = C.c_long
        L.orc_check_motions_batch.argtypes = [C.POINTER(Env), vp, vp, C.c_long, C.c_double, C.c_uint, vp, vp,
                                              C.c_int]
        L.orc_validate_states_batch.argtypes = [C.POINTER(Env), vp, C.c_long, vp, C.c_int]
        L.orc_knn_batch.argtypes = [C.c_int, vp, C.c_long, C.c_int, vp, C.c_long, C.c_int, vp, vp, C.c_int]
        L.orc_rrt_connect.argtypes = [C.POINTER(Env), vp, vp, C.c_uint32, C.c_double, C.c_double, C.c_uint,
                                      C.c_int, vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_long)]
        L.orc_rrt_connect_ex.argtypes = [C.POINTER(Env), vp, C.c_int, vp, C.c_int, C.c_uint32, C.c_double,
                                         C.c_double, C.c_uint, C.c_int, C.c_int, vp, C.c_int, C.POINTER(C.c_int),
                                         C.POINTER(C.c_int), C.POINTER(C.c_long)]
        L.orc_gnat_build.argtypes = [C.c_int, vp, C.c_int64, C.c_int, C.c_uint32]
        L.orc_gnat_build.restype = vp
        L.orc_gnat_free.argtypes = [vp]
        L.orc_gnat_knn.argtypes = [vp, vp, C.c_int64, C.c_int, vp, vp, C.c_int]
        L.orc_gnat_radius.argtypes = [vp, vp, C.c_int64, C.c_double, vp, vp, C.c_int64, C.c_int]
        L.orc_gnat_radius.restype = C.c_int64
        _lib = L
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(vp)


class OracleEnv:
    def __init__(self, spheres, boxes):
        self.s = np.ascontiguousarray(spheres, np.float32).reshape(-1, 4)
        self.b = np.ascontiguousarray(boxes, np.float32).reshape(-1, 6)
        self.c = Env(self.s.shape[0], _p(self.s), self.b.shape[0], _p(self.b))

    def ref(self):
        return C.byref(self.c)


def sincosf(x: float):
    s, c = C.c_float(), C.c_float()
    lib().orc_sincosf(x, C.byref(s), C.byref(c))
    return s.value, c.value


def acosf(x: float) -> float:
    return lib().orc_acosf(x)


def arm_spheres(q) -> np.ndarray:
    qa = np.ascontiguousarray(q, np.float32)
    out = np.zeros((28, 3), np.float32)
    lib().orc_arm_spheres(_p(qa), _p(out))
    return out


def state_valid(env: OracleEnv, q) -> bool:
    qa = np.ascontiguousarray(q, np.float32)
    return bool(lib().orc_state_valid(env.ref(), _p(qa), None))


def validate_states(env: OracleEnv, q_soa: np.ndarray) -> np.ndarray:
    qa = np.ascontiguousarray(q_soa, np.float32)
    n = qa.shape[1]
    out = np.zeros(n, np.uint8)
    lib().orc_validate_states(env.ref(), _p(qa), n, _p(out))
    return out


def max_extent(lo, hi) -> float:
    lo = np.ascontiguousarray(lo, np.float64)
    hi = np.ascontiguousarray(hi, np.float64)
    return lib().orc_max_extent(_p(lo), _p(hi), lo.shape[0])


def longest_valid_segment(max_ext: float, fraction: float) -> float:
    return lib().orc_longest_valid_segment(max_ext, fraction)


def valid_segment_count(a, b, lvs, factor=1) -> int:
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return lib().orc_valid_segment_count(_p(a), _p(b), a.shape[0], lvs, factor)


def bfs_order(nd: int) -> np.ndarray:
    out = np.zeros(max(nd, 1), np.int32)
    n = lib().orc_bfs_order(nd, _p(out))
    return out[:n].copy()


def check_motion(env: OracleEnv, a, b, lvs, factor=1):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    cnt = C.c_int()
    v = lib().orc_check_motion(env.ref(), _p(a), _p(b), lvs, factor, C.byref(cnt))
    return bool(v), cnt.value


def distance(a, b) -> float:
    """RealVectorStateSpace::distance (FP64)."""
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    return lib().orc_distance_d(_p(a), _p(b), a.shape[0])


def interpolate(a, b, t: float) -> np.ndarray:
    """RealVectorStateSpace::interpolate (FP64)."""
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    out = np.zeros_like(a)
    lib().orc_interpolate_d(_p(a), _p(b), float(t), a.shape[0], _p(out))
    return out


def check_states(env: OracleEnv, states):
    """SpaceInformation::checkMotion(states, count) in the reference's visit order -> (valid, checks)."""
    q = np.ascontiguousarray(np.asarray(states, np.float64).reshape(-1, 7))
    cnt = C.c_int()
    v = lib().orc_check_states(env.ref(), _p(q), q.shape[0], C.byref(cnt))
    return bool(v), cnt.value


def check_states_first_invalid(env: OracleEnv, states):
    """checkMotion(states, count, firstInvalidStateIndex) -> (valid, first invalid index or -1)."""
    q = np.ascontiguousarray(np.asarray(states, np.float64).reshape(-1, 7))
    f = C.c_int(-1)
    v = lib().orc_check_states_first_invalid(env.ref(), _p(q), q.shape[0], C.byref(f))
    return bool(v), f.value


def check_motion_last_valid(env: OracleEnv, a, b, lvs, factor=1, last_t=float("nan")):
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    cnt = C.c_int()
    t = C.c_double(last_t)
    v = lib().orc_check_motion_last_valid(env.ref(), _p(a), _p(b), lvs, factor, C.byref(t), C.byref(cnt))
    return bool(v), t.value, cnt.value


def check_motions(env: OracleEnv, a_rows: np.ndarray, b_rows: np.ndarray, lvs, factor=1, threads=1):
    """Row-major [n][7] endpoints -> (valid u8[n], checked i32[n])."""
    a = np.ascontiguousarray(a_rows, np.float64)
    b = np.ascontiguousarray(b_rows, np.float64)
    n = a.shape[0]
    out = np.zeros(n, np.uint8)
    chk = np.zeros(n, np.int32)
    lib().orc_check_motions_batch(env.ref(), _p(a), _p(b), n, lvs, factor, _p(out), _p(chk), threads)
    return out, chk


def rrt_connect(env: OracleEnv, start, goal, seed: int, max_distance: float, lvs: float, factor: int = 1,
                max_iters: int = 5000, cap: int = 4096):
    """RRT-Connect restatement -> (solved, iterations, path[n][7], tree_nodes)."""
    s = np.ascontiguousarray(start, np.float64)
    g = np.ascontiguousarray(goal, np.float64)
    path = np.zeros((cap, 7), np.float64)
    n, it, nodes = C.c_int(), C.c_int(), C.c_long()
    ok = lib().orc_rrt_connect(env.ref(), _p(s), _p(g), seed, max_distance, lvs, factor, max_iters, _p(path), cap,
                               C.byref(n), C.byref(it), C.byref(nodes))
    return bool(ok), it.value, path[:min(n.value, cap)].copy(), nodes.value


PLAN_FAILED, PLAN_SOLVED, PLAN_INVALID_START, PLAN_INVALID_GOAL = 0, 1, 2, 3


def rrt_connect_ex(env: OracleEnv, starts, goals, seed: int, max_distance: float, lvs: float, factor: int = 1,
                   max_iters: int = 5000, intermediate: bool = False, cap: int = 1 << 16):
    """RRT-Connect restatement with start and goal lists ([n][7]) and
    addIntermediateStates -> (status PLAN_*, iterations, path[n][7], tree_nodes)."""
    s = np.ascontiguousarray(np.reshape(starts, (-1, 7)), np.float64)
    g = np.ascontiguousarray(np.reshape(goals, (-1, 7)), np.float64)
    path = np.zeros((cap, 7), np.float64)
    n, it, nodes = C.c_int(), C.c_int(), C.c_long()
    st = lib().orc_rrt_connect_ex(env.ref(), _p(s), s.shape[0], _p(g), g.shape[0], seed, max_distance, lvs, factor,
                                  max_iters, int(bool(intermediate)), _p(path), cap, C.byref(n), C.byref(it),
                                  C.byref(nodes))
    return st, it.value, path[:min(n.value, cap)].copy(), nodes.value


def validate_states_mt(env: OracleEnv, q_soa: np.ndarray, threads: int) -> np.ndarray:
    qa = np.ascontiguousarray(q_soa, np.float32)
    n = qa.shape[1]
    out = np.zeros(n, np.uint8)
    lib().orc_validate_states_batch(env.ref(), _p(qa), n, _p(out), threads)
    return out


def knn_mt(metric: int, data: np.ndarray, queries: np.ndarray, k: int, threads: int):
    data = np.ascontiguousarray(data, np.float32)
    queries = np.ascontiguousarray(queries, np.float32)
    idx = np.zeros((queries.shape[0], k), np.int32)
    dist = np.zeros((queries.shape[0], k), np.float32)
    lib().orc_knn_batch(metric, _p(data), data.shape[0], data.shape[1], _p(queries), queries.shape[0], k, _p(idx),
                        _p(dist), threads)
    return idx, dist


class Mt19937:
    def __init__(self, seed: int):
        self.g = MT()
        lib().orc_mt_seed(C.byref(self.g), seed)

    def next(self) -> int:
        return lib().orc_mt_next(C.byref(self.g))

    def canonical(self) -> float:
        return lib().orc_canonical(C.byref(self.g))

    def uniform_real(self, lo, hi) -> float:
        return lib().orc_uniform_real(C.byref(self.g), lo, hi)

    def sample_uniform(self, lo, hi, n) -> np.ndarray:
        lo = np.ascontiguousarray(lo, np.float64)
        hi = np.ascontiguousarray(hi, np.float64)
        out = np.zeros((n, lo.shape[0]), np.float64)
        lib().orc_sample_uniform(C.byref(self.g), _p(lo), _p(hi), lo.shape[0], n, _p(out))
        return out

    def sample_valid(self, env: "OracleEnv", n_calls: int, attempts: int = 100, lo=None, hi=None):
        """UniformValidStateSampler::sample x n_calls -> (states [n][7], found u8 [n], draws i32 [n])."""
        lo = np.ascontiguousarray(np.full(7, -np.pi) if lo is None else lo, np.float64)
        hi = np.ascontiguousarray(np.full(7, np.pi) if hi is None else hi, np.float64)
        out = np.zeros((n_calls, 7), np.float64)
        found = np.zeros(n_calls, np.uint8)
        draws = np.zeros(n_calls, np.int32)
        lib().orc_sample_valid(env.ref(), C.byref(self.g), _p(lo), _p(hi), n_calls, attempts, _p(out), _p(found),
                               _p(draws))
        return out, found, draws

    def quaternion(self) -> np.ndarray:
        q = np.zeros(4, np.float64)
        lib().orc_quaternion(C.byref(self.g), _p(q))
        return q


def l2_dist(e, q) -> float:
    e = np.ascontiguousarray(e, np.float32)
    q = np.ascontiguousarray(q, np.float32)
    return lib().orc_l2_dist(_p(e), _p(q), e.shape[0])


def se3_dist(e, q) -> float:
    e = np.ascontiguousarray(e, np.float32)
    q = np.ascontiguousarray(q, np.float32)
    return lib().orc_se3_dist(_p(e), _p(q))


def knn(metric: int, data: np.ndarray, queries: np.ndarray, k: int):
    data = np.ascontiguousarray(data, np.float32)
    queries = np.ascontiguousarray(queries, np.float32)
    n, dim = data.shape
    nq = queries.shape[0]
    idx = np.zeros((nq, k), np.int32)
    dist = np.zeros((nq, k), np.float32)
    for j in range(nq):
        lib().orc_knn(metric, _p(data), n, dim, _p(queries[j]), k, _p(idx[j]), _p(dist[j]))
    return idx, dist


def radius(metric: int, data: np.ndarray, queries: np.ndarray, r: float):
    data = np.ascontiguousarray(data, np.float32)
    queries = np.ascontiguousarray(queries, np.float32)
    n, dim = data.shape
    res = []
    cap = 4096
    for j in range(queries.shape[0]):
        while True:
            idx = np.zeros(cap, np.int32)
            dist = np.zeros(cap, np.float32)
            cnt = lib().orc_radius(metric, _p(data), n, dim, _p(queries[j]), r, _p(idx), _p(dist), cap)
            if cnt <= cap:
                break
            cap = int(cnt)
        res.append((idx[:cnt].copy(), dist[:cnt].copy()))
    return res


def nearest(metric: int, data: np.ndarray, query) -> int:
    data = np.ascontiguousarray(data, np.float32)
    q = np.ascontiguousarray(query, np.float32)
    return lib().orc_nearest(metric, _p(data), data.shape[0], data.shape[1], _p(q))


def ref_linear(metric: int, data: np.ndarray, queries: np.ndarray, k: int, r: float, workdir: Path):
    """Runs the reference NearestNeighborsLinear driver (FP64 distances)."""
    if not REF_LINEAR.exists():
        raise FileNotFoundError(REF_LINEAR)
    workdir = Path(workdir)
    d = workdir / "data.f32"
    q = workdir / "queries.f32"
    o = workdir / "out.txt"
    np.ascontiguousarray(data, np.float32).tofile(d)
    np.ascontiguousarray(queries, np.float32).tofile(q)
    subprocess.run([str(REF_LINEAR), str(metric), str(data.shape[1]), str(d), str(data.shape[0]), str(q),
                    str(queries.shape[0]), str(k), repr(float(r)), str(o)], check=True)
    knn_res, rad_res, nearest_res = [], [], []
    for line in o.read_text().splitlines():
        parts = line.split()
        if parts[0] == "K":
            vals = parts[1:]
            knn_res.append((np.array(vals[0::2], np.int64), np.array(vals[1::2], np.float64)))
        elif parts[0] == "R":
            vals = parts[2:]
            rad_res.append((np.array(vals[0::2], np.int64), np.array(vals[1::2], np.float64)))
        else:
            nearest_res.append(int(parts[1]))
    return knn_res, rad_res, nearest_res


def disjoint_set_labels(n: int, edges: np.ndarray, valid: np.ndarray) -> np.ndarray:
    """PRM::uniteComponents over the valid edges (geometric/planners/prm/src/PRM.cpp:597, 606-609:
    boost::disjoint_sets union_set), restated as union by smaller root with path halving; the label
    of a component is its smallest element."""
    parent = list(range(n))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    for (u, v), ok in zip(edges.tolist(), valid.tolist()):
        if ok:
            ru, rv = find(u), find(v)
            if ru != rv:
                parent[max(ru, rv)] = min(ru, rv)
    return np.array([find(i) for i in range(n)], np.int32)


def kstar_k(n_milestones: int, dim: int = 7) -> int:
    """KStarStrategy's k (geometric/planners/prm/ConnectionStrategy.h:137-150): ceil((e + e/d) log n),
    capped at 63 (the device k-NN bound)."""
    import math

    kc = math.e + math.e / dim
    return max(1, min(63, math.ceil(kc * math.log(n_milestones))))


def prm_batch(env: OracleEnv, seed: int, n_samples: int, k: int, lvs: float, factor: int = 1, n_streams: int = 1,
              threads: int = 1, kstar: bool = False, extra=None):
    """Batch PRM roadmap: the steps of PRM::growRoadmap / addMilestone (geometric/planners/prm/src/
    PRM.cpp:364-385, 570-604) with KStrategy (geometric/planners/prm/ConnectionStrategy.h:56-81) over a
    fixed sample set.  Samples: stream s = mt19937(seed + s), n_samples rounded up to a multiple of
    n_streams, stream-major; milestones = valid samples in order; candidates = undirected union of the
    k nearest other milestones of each milestone (FP32 L2, ties to the lower index); each candidate
    (u < v) checked from u to v (PRM.cpp:590 checks from the older milestone to the new one).
    Returns a dict of numpy arrays."""
    per = -(-n_samples // n_streams)
    lo, hi = np.full(7, -np.pi), np.full(7, np.pi)
    samples = np.concatenate([Mt19937(seed + s).sample_uniform(lo, hi, per) for s in range(n_streams)])
    if extra is not None:  # query states placed before the samples (milestones 0, 1, ... when valid)
        samples = np.concatenate([np.asarray(extra, np.float64).reshape(-1, 7), samples])
    valid = validate_states_mt(env, samples.T.astype(np.float32), threads).astype(bool)
    sidx = np.nonzero(valid)[0].astype(np.int32)
    nodes = samples[sidx]
    nv = nodes.shape[0]
    if kstar and nv >= 1:
        k = kstar_k(nv)
    out = dict(samples=samples, valid=valid, sample_index=sidx, milestones=nodes, k=k)
    if nv < 2:
        out.update(edges=np.zeros((0, 2), np.int32), edge_valid=np.zeros(0, bool), component=np.arange(nv,
                   dtype=np.int32))
        return out
    kk = min(k + 1, nv)
    f32 = nodes.astype(np.float32)
    idx, _ = knn_mt(0, f32, f32, kk, threads)
    keys = set()
    for i in range(nv):
        taken = 0
        for j in idx[i].tolist():
            if taken < k and j >= 0 and j != i:
                keys.add((min(i, j), max(i, j)))
                taken += 1
    edges = np.array(sorted(keys), np.int32).reshape(-1, 2)
    ok, _ = check_motions(env, nodes[edges[:, 0]], nodes[edges[:, 1]], lvs, factor, threads)
    out.update(edges=edges, edge_valid=ok.astype(bool), component=disjoint_set_labels(nv, edges, ok.astype(bool)))
    return out


class Gnat:
    """NearestNeighborsGNAT restatement (oracle/gnat.cpp) over FP32 points with
    FP64 distances (metric 0: L2, 1: SE3)."""

    def __init__(self, metric: int, data: np.ndarray, seed: int = 1):
        self.data = np.ascontiguousarray(data, np.float32)
        self.metric, self.dim = metric, self.data.shape[1]
        self.h = lib().orc_gnat_build(metric, _p(self.data), self.data.shape[0], self.dim, seed)
        if not self.h:
            raise ValueError("GNAT build: empty set or bad dimension")

    def knn(self, queries: np.ndarray, k: int, threads: int = 1):
        q = np.ascontiguousarray(queries, np.float32)
        idx = np.zeros((q.shape[0], k), np.int32)
        dist = np.zeros((q.shape[0], k), np.float64)
        lib().orc_gnat_knn(self.h, _p(q), q.shape[0], k, _p(idx), _p(dist), threads)
        return idx, dist

    def radius(self, queries: np.ndarray, r: float, threads: int = 1):
        """-> (offsets [nq+1], ascending index lists)"""
        q = np.ascontiguousarray(queries, np.float32)
        counts = np.zeros(q.shape[0], np.int64)
        total = lib().orc_gnat_radius(self.h, _p(q), q.shape[0], r, _p(counts), None, 0, threads)
        idx = np.zeros(max(total, 1), np.int32)
        lib().orc_gnat_radius(self.h, _p(q), q.shape[0], r, _p(counts), _p(idx), total, threads)
        return np.concatenate([[0], np.cumsum(counts)]), idx[:total]

    def close(self):
        if getattr(self, "h", None):
            lib().orc_gnat_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def roadmap_shortest_path(nv: int, edges: np.ndarray, valid: np.ndarray, nodes: np.ndarray):
    """Shortest path from milestone 0 to milestone 1 over the valid edges, FP64
    state distances as weights (the weight PRM::constructSolution's A* uses,
    geometric/planners/prm/src/PRM.cpp:696-727); nodes popped in (distance,
    index) order, relaxed only on a strictly shorter distance.  Node list, or
    [] when 1 is unreachable."""
    import heapq
    import math

    adj = [[] for _ in range(nv)]
    for (u, v), ok in zip(edges.tolist(), valid.tolist()):
        if ok:
            adj[u].append(v)
            adj[v].append(u)
    for a in adj:
        a.sort()
    dist = [math.inf] * nv
    prev = [-1] * nv
    dist[0] = 0.0
    pq = [(0.0, 0)]
    while pq:
        d, u = heapq.heappop(pq)
        if d > dist[u]:
            continue
        if u == 1:
            break
        for v in adj[u]:
            s = 0.0
            for kk in range(7):
                diff = float(nodes[u, kk]) - float(nodes[v, kk])
                s += diff * diff
            nd = d + math.sqrt(s)
            if nd < dist[v]:
                dist[v] = nd
                prev[v] = u
                heapq.heappush(pq, (nd, v))
    if nv < 2 or not math.isfinite(dist[1]):
        return []
    path, v = [], 1
    while v >= 0:
        path.append(v)
        v = prev[v]
    return path[::-1]


def prm_plan(env: OracleEnv, start, goal, seed: int, n_samples: int, k: int, lvs: float, factor: int = 1,
             n_streams: int = 64, threads: int = 1):
    """A served PRM request: start and goal as milestones 0 and 1 of the batch
    roadmap, then the shortest path.  -> (solved, path [n][7], milestones, candidate edges)."""
    r = prm_batch(env, seed, n_samples, k, lvs, factor, n_streams, threads, extra=np.stack([start, goal]))
    idx = r["sample_index"]
    nv = len(idx)
    if nv < 2 or idx[0] != 0 or idx[1] != 1:
        return False, np.zeros((0, 7)), nv, len(r["edges"])
    p = roadmap_shortest_path(nv, r["edges"], r["edge_valid"], r["milestones"])
    return bool(p), r["milestones"][p] if p else np.zeros((0, 7)), nv, len(r["edges"])
