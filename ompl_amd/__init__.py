"""ompl_amd: MI355X (gfx950) implementation of OMPL's data-parallel planning
inner loop -- batched state validity (FP32 arm FK + sphere/box collision),
DiscreteMotionValidator edge checks, uniform state sampling and exact
nearest-neighbour queries -- behind a C-ABI (include/ompl_amd.h).

This module is a thin Python host layer over that C-ABI, used by the tests and
the benchmark.  Device buffers are torch tensors on the current HIP stream
(torch is only the allocator / stream provider here); every computation runs
in libompl_amd.so and raises if the library is not built.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from ._lib import OmplAmdError, Tuning, call, check, load  # noqa: F401

ARM_DOF = 7
ARM_SPHERES = 28
METRIC_L2 = 0
METRIC_SE3 = 1


def _torch():
    import torch

    return torch


def _stream(stream):
    """The HIP stream to launch on: an explicit handle, else torch's current
    stream (its default stream is HIP's null stream, handle 0)."""
    if stream is not None:
        return C.c_void_p(stream)
    torch = _torch()
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


def init(device: int = 0) -> None:
    call("ompl_amd_init", device)


def capabilities() -> dict:
    """Limits of the library and the device (initialises device 0 if needed)."""
    from ._lib import Caps

    c = Caps()
    call("ompl_amd_capabilities", C.byref(c))
    d = {k: getattr(c, k) for k, _ in c._fields_}
    d["device_name"] = c.device_name.decode()
    return d


def get_tuning() -> dict:
    t = Tuning()
    call("ompl_amd_get_tuning", C.byref(t))
    return t.as_dict()


def set_tuning(**kw) -> dict:
    t = Tuning()
    call("ompl_amd_get_tuning", C.byref(t))
    for k, v in kw.items():
        if not hasattr(t, k):
            raise KeyError(k)
        setattr(t, k, int(v))
    call("ompl_amd_set_tuning", C.byref(t))
    return t.as_dict()


class profile:
    """Per-kernel HIP-event timing inside the library (launch stream)."""

    @staticmethod
    def enable(on: bool = True) -> None:
        call("ompl_amd_profile_enable", 1 if on else 0)

    @staticmethod
    def reset() -> None:
        call("ompl_amd_profile_reset")

    @staticmethod
    def query(kernel: str):
        n = C.c_int64()
        ms = C.c_double()
        call("ompl_amd_profile_query", kernel.encode(), C.byref(n), C.byref(ms))
        return n.value, ms.value


def motion_work(env: "Env", a_soa, b_soa, lvs: float, factor: int = 1) -> dict:
    out = (C.c_int64 * 3)()
    call("ompl_amd_motion_work", env.handle, _ptr(a_soa), _ptr(b_soa), a_soa.shape[1], float(lvs), int(factor),
         out)
    return {"states": out[0], "sphere_pairs": out[1], "box_pairs": out[2]}


def validity_work(env: "Env", q_soa) -> dict:
    out = (C.c_int64 * 3)()
    call("ompl_amd_validity_work", env.handle, _ptr(q_soa), q_soa.shape[1], out)
    return {"states": out[0], "sphere_pairs": out[1], "box_pairs": out[2]}


# FP32 operations of the checker's formulas (see arm_device.hpp):
#   per joint: sincos 30 (reduction 8, polynomials 20, quadrant 2), rotation
#   update 18, four sphere centres 24, joint position 6  -> 78; 7 joints -> 546
#   sphere/sphere pair: 3 sub + 3 mul + 2 add + 1 cmp = 9
#   sphere/box pair: 3 max + 3 min + 3 sub + 3 mul + 2 add + 1 cmp = 15
FK_FLOPS = 546
SPHERE_PAIR_FLOPS = 9
BOX_PAIR_FLOPS = 15


def work_flops(w: dict) -> int:
    return w["states"] * FK_FLOPS + w["sphere_pairs"] * SPHERE_PAIR_FLOPS + w["box_pairs"] * BOX_PAIR_FLOPS


def autotune_motions(env: "Env", a_soa, b_soa, lvs: float, factor: int = 1, reps: int = 20):
    """Runs the library's launch-shape autotuner on a device-resident batch and
    installs the fastest shape; returns (tuning dict, microseconds per batch)."""
    t = Tuning()
    us = C.c_double()
    call("ompl_amd_autotune_motions", env.handle, _ptr(a_soa), _ptr(b_soa), a_soa.shape[1], float(lvs),
         int(factor), int(reps), C.byref(t), C.byref(us))
    return t.as_dict(), us.value


def autotune_motions_streams(env: "Env", a_soa, b_soa, lvs: float, streams, factor: int = 1, steps: int = 200):
    """Runs the library's autotuner for batches issued round-robin on
    `streams` (torch streams whose batches overlap) and installs the fastest
    shape; returns (tuning dict, microseconds per batch).

    The shape that is fastest for one batch at a time is not the fastest when
    batches of several streams share the device: fewer persistent walk waves
    leave compute units to the other streams' s2 passes.  The library times
    each shape of its grid (s2 pass x walk shape x walk waves per CU, the
    chunk-parallel walk) over `steps` batches on the streams, and the three
    fastest again over 2 x `steps`.  Results do not depend on the shape.  An
    environment with its own shapes is refused (they would take the place of
    every candidate)."""
    streams = list(streams)
    handles = (C.c_void_p * max(1, len(streams)))(*[s.cuda_stream for s in streams])
    t = Tuning()
    us = C.c_double()
    call("ompl_amd_autotune_motions_streams", env.handle, _ptr(a_soa), _ptr(b_soa), a_soa.shape[1], float(lvs),
         int(factor), handles, len(streams), int(steps), C.byref(t), C.byref(us))
    return t.as_dict(), us.value


def longest_valid_segment(lo, hi, resolution: float) -> float:
    """StateSpace::setup: maxExtent * fraction (base/src/StateSpace.cpp:237-241),
    with fraction = resolution / maxExtent as SpaceInformation::
    setStateValidityCheckingResolution would be given for an absolute step."""
    e = 0.0
    for a, b in zip(lo, hi):
        d = b - a
        e += d * d
    ext = math.sqrt(e)
    return ext * (resolution / ext)


class Env:
    """Device-resident obstacle set (spheres (x,y,z,R); boxes (cx,cy,cz,hx,hy,hz))."""

    def __init__(self, spheres=None, boxes=None):
        sph = np.ascontiguousarray(np.zeros((0, 4), np.float32) if spheres is None else spheres, np.float32)
        box = np.ascontiguousarray(np.zeros((0, 6), np.float32) if boxes is None else boxes, np.float32)
        assert sph.ndim == 2 and sph.shape[1] == 4 and box.ndim == 2 and box.shape[1] == 6
        self.spheres, self.boxes = sph, box
        h = C.c_void_p()
        call("ompl_amd_env_create", sph.ctypes.data_as(C.c_void_p), sph.shape[0],
             box.ctypes.data_as(C.c_void_p), box.shape[0], C.byref(h))
        self.handle = h

    @property
    def nbytes(self) -> int:
        b = C.c_size_t()
        call("ompl_amd_env_bytes", self.handle, C.byref(b))
        return b.value

    def set_tuning(self, tuning: dict | None = None, **kw) -> dict:
        """Launch shapes of this environment's batched checks: the fields of
        `tuning` and `kw` over the shapes now in force for it; set_tuning(None)
        with no keywords makes it follow the library's tuning again."""
        if tuning is None and not kw:
            call("ompl_amd_env_set_tuning", self.handle, None)
            return self.tuning()
        t = Tuning()
        call("ompl_amd_env_get_tuning", self.handle, C.byref(t), None)
        for k, v in dict(tuning or {}, **kw).items():
            if not hasattr(t, k):
                raise KeyError(k)
            setattr(t, k, int(v))
        call("ompl_amd_env_set_tuning", self.handle, C.byref(t))
        return t.as_dict()

    def tuning(self) -> dict:
        """The launch shapes in force for this environment, with "own": True
        when they are its own rather than the library's."""
        t = Tuning()
        own = C.c_int()
        call("ompl_amd_env_get_tuning", self.handle, C.byref(t), C.byref(own))
        return dict(t.as_dict(), own=bool(own.value))

    def autotune(self, a_soa, b_soa, lvs: float, factor: int = 1, reps: int = 20):
        """Runs the motion autotuner on a device-resident batch of this
        environment and keeps the winner as the environment's own shapes (the
        library's tuning is left as it was); returns (tuning, us per batch)."""
        saved = get_tuning()
        try:
            t, us = autotune_motions(self, a_soa, b_soa, lvs, factor, reps)
        finally:
            set_tuning(**saved)
        self.set_tuning(t)
        return t, us

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            call("ompl_amd_env_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def validate_states(env: Env, q_soa, out=None, stream=None):
    """Batched StateValidityChecker::isValid; q_soa: device [7, n] float32 or float64."""
    torch = _torch()
    assert q_soa.is_cuda and q_soa.dim() == 2 and q_soa.shape[0] == ARM_DOF and q_soa.is_contiguous()
    n = q_soa.shape[1]
    if out is None:
        out = torch.empty(n, dtype=torch.uint8, device=q_soa.device)
    fn = "ompl_amd_validate_states_dev" if q_soa.dtype == torch.float32 else "ompl_amd_validate_states_f64_dev"
    if q_soa.dtype not in (torch.float32, torch.float64):
        raise TypeError("states must be float32 or float64")
    call(fn, env.handle, _ptr(q_soa), n, _ptr(out), _stream(stream))
    return out


def arm_spheres(q_soa, stream=None):
    torch = _torch()
    n = q_soa.shape[1]
    out = torch.empty((n, ARM_SPHERES, 3), dtype=torch.float32, device=q_soa.device)
    call("ompl_amd_arm_spheres_dev", _ptr(q_soa), n, _ptr(out), _stream(stream))
    return out


def check_motions(env: Env, a_soa, b_soa, lvs: float, factor: int = 1, out=None, checked=None,
                  exact_count: bool = False, stream=None):
    """Batched DiscreteMotionValidator::checkMotion(s1, s2); a_soa/b_soa: device [7, n] float64."""
    torch = _torch()
    assert a_soa.dtype == torch.float64 and b_soa.dtype == torch.float64
    assert a_soa.shape == b_soa.shape and a_soa.shape[0] == ARM_DOF
    assert a_soa.is_contiguous() and b_soa.is_contiguous()
    n = a_soa.shape[1]
    if out is None:
        out = torch.empty(n, dtype=torch.uint8, device=a_soa.device)
    if exact_count and checked is None:
        checked = torch.empty(n, dtype=torch.int32, device=a_soa.device)
    call("ompl_amd_check_motions_dev", env.handle, _ptr(a_soa), _ptr(b_soa), n, float(lvs), int(factor),
         _ptr(out), _ptr(checked), _stream(stream))
    return (out, checked) if checked is not None else out


def check_motions_last_valid(env: Env, a_soa, b_soa, lvs: float, factor: int = 1, stream=None):
    torch = _torch()
    n = a_soa.shape[1]
    out = torch.empty(n, dtype=torch.uint8, device=a_soa.device)
    last_t = torch.full((n,), float("nan"), dtype=torch.float64, device=a_soa.device)
    call("ompl_amd_check_motions_last_valid_dev", env.handle, _ptr(a_soa), _ptr(b_soa), n, float(lvs),
         int(factor), _ptr(out), _ptr(last_t), _stream(stream))
    return out, last_t


def is_valid(env: Env, q) -> bool:
    """Plain single-call path: one launch per state."""
    qa = np.ascontiguousarray(q, np.float64)
    v = C.c_int()
    call("ompl_amd_is_valid", env.handle, qa.ctypes.data_as(C.c_void_p), C.byref(v))
    return bool(v.value)


def check_motion(env: Env, a, b, lvs: float, factor: int = 1) -> bool:
    """Plain single-call path: one launch per checkMotion(s1, s2)."""
    aa = np.ascontiguousarray(a, np.float64)
    bb = np.ascontiguousarray(b, np.float64)
    v = C.c_int()
    call("ompl_amd_check_motion", env.handle, aa.ctypes.data_as(C.c_void_p), bb.ctypes.data_as(C.c_void_p),
         float(lvs), int(factor), C.byref(v))
    return bool(v.value)


def check_motion_last_valid(env: Env, a, b, lvs: float, factor: int = 1, last_t: float = float("nan")):
    aa = np.ascontiguousarray(a, np.float64)
    bb = np.ascontiguousarray(b, np.float64)
    v = C.c_int()
    t = C.c_double(last_t)
    call("ompl_amd_check_motion_last_valid", env.handle, aa.ctypes.data_as(C.c_void_p),
         bb.ctypes.data_as(C.c_void_p), float(lvs), int(factor), C.byref(v), C.byref(t))
    return bool(v.value), t.value


class MotionGraph:
    """A motion batch captured in a HIP graph: launch() replays exactly the
    check_motions call it was made from (same device arrays, sizes and launch
    shape) on the given stream.  Replays of one graph must not overlap."""

    def __init__(self, env: Env, a_soa, b_soa, lvs: float, factor: int = 1, out=None, checked=None):
        torch = _torch()
        n = a_soa.shape[1]
        self.out = out if out is not None else torch.empty(n, dtype=torch.uint8, device=a_soa.device)
        self.checked = checked
        self._keep = (a_soa, b_soa)  # the graph reads these arrays on every replay
        h = C.c_void_p()
        call("ompl_amd_motion_graph_create", env.handle, _ptr(a_soa), _ptr(b_soa), n, float(lvs), int(factor),
             _ptr(self.out), _ptr(checked), C.byref(h))
        self.handle = h

    def launch(self, stream=None):
        call("ompl_amd_motion_graph_launch", self.handle, _stream(stream))
        return self.out

    @classmethod
    def group(cls, env: Env, batches, lvs: float, factor: int = 1, outs=None, checked=None):
        """Independent batches [(a_soa, b_soa), ...] captured in one graph,
        each on its own forked stream: launch() runs them concurrently.
        `outs` / `checked` are per-batch lists (created when None)."""
        torch = _torch()
        self = cls.__new__(cls)
        nb = len(batches)
        if outs is None:
            outs = [torch.empty(a.shape[1], dtype=torch.uint8, device=a.device) for a, _ in batches]
        self.out = outs
        self.checked = checked
        self._keep = (batches, outs, checked)
        vpn = C.c_void_p * nb
        a_arr = vpn(*[a.data_ptr() for a, _ in batches])
        b_arr = vpn(*[b.data_ptr() for _, b in batches])
        n_arr = (C.c_int64 * nb)(*[a.shape[1] for a, _ in batches])
        o_arr = vpn(*[o.data_ptr() for o in outs])
        c_arr = vpn(*[c.data_ptr() for c in checked]) if checked is not None else None
        h = C.c_void_p()
        call("ompl_amd_motion_graph_create_group", env.handle, nb, a_arr, b_arr, n_arr, float(lvs), int(factor),
             o_arr, c_arr, C.byref(h))
        self.handle = h
        return self

    def close(self):
        if getattr(self, "handle", None) is not None and self.handle.value:
            call("ompl_amd_motion_graph_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def check_state_lists(env: Env, q_soa, offsets, stream=None):
    """Batched SpaceInformation::checkMotion(states, count[, firstInvalidStateIndex])
    (base/src/SpaceInformation.cpp:288-341) over lists of precomputed states.
    q_soa: device [7, n_states] float64; offsets: device int64 [n_lists + 1].
    Returns (valid u8 [n_lists], first_invalid i32 [n_lists], -1 when valid)."""
    torch = _torch()
    assert q_soa.dtype == torch.float64 and q_soa.shape[0] == ARM_DOF and q_soa.is_contiguous()
    assert offsets.dtype == torch.int64 and offsets.is_contiguous()
    n_lists = offsets.shape[0] - 1
    out = torch.empty(max(n_lists, 0), dtype=torch.uint8, device=q_soa.device)
    first = torch.empty(max(n_lists, 0), dtype=torch.int32, device=q_soa.device)
    call("ompl_amd_check_state_lists_dev", env.handle, _ptr(q_soa), q_soa.shape[1], _ptr(offsets), n_lists,
         _ptr(out), _ptr(first), _stream(stream))
    return out, first


def check_states(env: Env, states) -> tuple[bool, int]:
    """Plain single-call path of SpaceInformation::checkMotion(states, count,
    firstInvalidStateIndex): returns (valid, first invalid index or -1)."""
    qa = np.ascontiguousarray(np.asarray(states, np.float64).reshape(-1, ARM_DOF))
    v, f = C.c_int(), C.c_int(-1)
    call("ompl_amd_check_states", env.handle, qa.ctypes.data_as(C.c_void_p), qa.shape[0], C.byref(v), C.byref(f))
    return bool(v.value), f.value


# ---------------------------------------------------------------------------
# synthetic workloads (host, std::mt19937 with ompl::RNG semantics)
# ---------------------------------------------------------------------------
class workload:
    @staticmethod
    def obstacles(seed, n_spheres, n_boxes, r=(0.05, 0.15), h=(0.05, 0.2), min_base_dist=0.2):
        sph = np.zeros((n_spheres, 4), np.float32)
        box = np.zeros((n_boxes, 6), np.float32)
        call("ompl_amd_workload_obstacles", seed, n_spheres, n_boxes, r[0], r[1], h[0], h[1], min_base_dist,
             sph.ctypes.data_as(C.c_void_p), box.ctypes.data_as(C.c_void_p))
        return sph, box

    @staticmethod
    def configs(seed, n):
        out = np.zeros((n, ARM_DOF), np.float64)
        call("ompl_amd_workload_configs", seed, n, out.ctypes.data_as(C.c_void_p))
        return out

    @staticmethod
    def edges(seed, n, mu=-1.0, sigma=1.0, len_lo=0.05, len_hi=5.0):
        a = np.zeros((n, ARM_DOF), np.float64)
        b = np.zeros((n, ARM_DOF), np.float64)
        call("ompl_amd_workload_edges", seed, n, mu, sigma, len_lo, len_hi, a.ctypes.data_as(C.c_void_p),
             b.ctypes.data_as(C.c_void_p))
        return a, b

    @staticmethod
    def se3(seed, n, lo=-10.0, hi=10.0):
        out = np.zeros((n, 7), np.float32)
        call("ompl_amd_workload_se3", seed, n, lo, hi, out.ctypes.data_as(C.c_void_p))
        return out


def rrtstar_radius(dim: int, measure: float, n: int, rewire_factor: float = 1.1,
                   max_distance: float = float("inf")) -> float:
    return load().ompl_amd_rrtstar_radius(dim, measure, rewire_factor, max_distance, n)
