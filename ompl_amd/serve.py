"""Python client of the planning server (continuous batching of RRT-Connect
requests with device-resident solve loops, include/ompl_amd.h) and of the
plain one-launch-per-call planner."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import call

DOF = 7
QUEUED, RUNNING, SOLVED, FAILED, DEADLINE, CANCELLED, REJECTED, TREE_FULL = range(8)
STATUS = {0: "queued", 1: "running", 2: "solved", 3: "failed", 4: "deadline", 5: "cancelled", 6: "rejected",
          7: "tree_full", 8: "invalid_start", 9: "invalid_goal"}


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_size_t), ("max_active", C.c_int), ("max_nodes", C.c_int),
                ("check_obstacle_parts", C.c_int), ("check_lanes_per_state", C.c_int), ("slice_us", C.c_int),
                ("block_threads", C.c_int)]


class PlanRequest(C.Structure):
    _fields_ = [
        ("env", C.c_void_p),
        ("start", C.c_double * DOF),
        ("goal", C.c_double * DOF),
        ("seed", C.c_uint32),
        ("max_iters", C.c_int),
        ("range", C.c_double),
        ("longest_valid_segment", C.c_double),
        ("factor", C.c_uint),
        ("deadline_ms", C.c_double),
        ("max_nodes", C.c_int),
        ("planner", C.c_int),
        ("prm_samples", C.c_int),
        ("prm_k", C.c_int),
        ("starts", C.c_void_p),
        ("n_starts", C.c_int),
        ("goals", C.c_void_p),
        ("n_goals", C.c_int),
        ("intermediate_states", C.c_int),
    ]


class PlanResult(C.Structure):
    _fields_ = [
        ("status", C.c_int),
        ("iterations", C.c_int),
        ("path_len", C.c_int),
        ("tree_nodes", C.c_int64),
        ("queued_ms", C.c_double),
        ("run_ms", C.c_double),
        ("latency_ms", C.c_double),
        ("motion_checks", C.c_int64),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["status_name"] = STATUS.get(self.status, "?")
        return d


RRT_CONNECT, PRM = 0, 1


def make_request(env, start, goal, seed, lvs, max_iters=5000, rng=0.0, factor=1, deadline_ms=0.0, max_nodes=0,
                 planner=RRT_CONNECT, prm_samples=0, prm_k=0, starts=None, goals=None, intermediate_states=False):
    """A plan request.  starts / goals ([n][7]) replace start / goal when
    given (the arrays are kept on the request until it is submitted)."""
    r = PlanRequest()
    r.intermediate_states = int(bool(intermediate_states))
    if starts is not None:
        a = np.ascontiguousarray(np.reshape(starts, (-1, DOF)), np.float64)
        r._starts = a
        r.starts = a.ctypes.data
        r.n_starts = a.shape[0]
        start = a[0] if start is None else start
    if goals is not None:
        a = np.ascontiguousarray(np.reshape(goals, (-1, DOF)), np.float64)
        r._goals = a
        r.goals = a.ctypes.data
        r.n_goals = a.shape[0]
        goal = a[0] if goal is None else goal
    r.planner = int(planner)
    r.prm_samples = int(prm_samples)
    r.prm_k = int(prm_k)
    r.env = env.handle.value if hasattr(env, "handle") else env
    for k in range(DOF):
        r.start[k] = float(start[k])
        r.goal[k] = float(goal[k])
    r.seed = int(seed)
    r.max_iters = int(max_iters)
    r.range = float(rng)
    r.longest_valid_segment = float(lvs)
    r.factor = int(factor)
    r.deadline_ms = float(deadline_ms)
    r.max_nodes = int(max_nodes)
    return r


def footprint(max_nodes: int) -> int:
    """Bytes of one request's device region for max_nodes nodes per tree."""
    from ._lib import load

    return int(load().ompl_amd_server_footprint(int(max_nodes)))


class Server:
    def __init__(self, hbm_budget_bytes=1 << 30, max_active=256, max_nodes=8192, check_obstacle_parts=0,
                 check_lanes_per_state=0, slice_us=0, block_threads=0):
        cfg = ServerConfig(hbm_budget_bytes, max_active, max_nodes, check_obstacle_parts, check_lanes_per_state,
                           slice_us, block_threads)
        h = C.c_void_p()
        call("ompl_amd_server_create", C.byref(cfg), C.byref(h))
        self.handle = h

    def submit(self, req: PlanRequest) -> int:
        rid = C.c_int64()
        call("ompl_amd_server_submit", self.handle, C.byref(req), C.byref(rid))
        return rid.value

    def poll(self, rid: int) -> PlanResult:
        r = PlanResult()
        call("ompl_amd_server_poll", self.handle, rid, C.byref(r))
        return r

    def wait(self, rid: int, timeout_ms: float = -1.0) -> PlanResult:
        r = PlanResult()
        call("ompl_amd_server_wait", self.handle, rid, float(timeout_ms), C.byref(r))
        return r

    def path(self, rid: int) -> np.ndarray:
        n = C.c_int()
        call("ompl_amd_server_path", self.handle, rid, None, 0, C.byref(n))
        out = np.zeros((max(n.value, 1), DOF), np.float64)
        call("ompl_amd_server_path", self.handle, rid, out.ctypes.data_as(C.c_void_p), n.value, C.byref(n))
        return out[:n.value]

    def cancel(self, rid: int) -> None:
        call("ompl_amd_server_cancel", self.handle, rid)

    def release(self, rid: int) -> None:
        call("ompl_amd_server_release", self.handle, rid)

    def stats(self) -> dict:
        out = (C.c_int64 * 6)()
        call("ompl_amd_server_stats", self.handle, out)
        return dict(zip(("slices", "steps", "active", "queued", "used_bytes", "budget_bytes"), list(out)))

    def phase_us(self) -> dict:
        """Device time of the growTree steps so far (summed over requests), us."""
        out = (C.c_double * 4)()
        call("ompl_amd_server_phase_us", self.handle, out)
        return dict(zip(("sample", "nearest", "check", "step"), list(out)))

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            call("ompl_amd_server_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def plan_plain(req: PlanRequest, cap: int = 4096):
    """RRT-Connect with one GPU launch + sync per isValid / checkMotion call."""
    res = PlanResult()
    path = np.zeros((cap, DOF), np.float64)
    n = C.c_int()
    call("ompl_amd_plan_plain", C.byref(req), C.byref(res), path.ctypes.data_as(C.c_void_p), cap, C.byref(n))
    return res, path[:min(n.value, cap)]
