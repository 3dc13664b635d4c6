"""ctypes binding of libompl_amd.so (the C-ABI declared in include/ompl_amd.h).

The library is loaded from the package directory only; there is no fallback:
if libompl_amd.so is missing or does not load, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libompl_amd.so"

_lib = None
# the library's ompl_amd_version() these signatures were written for (OMPL_AMD_VERSION)
ABI_VERSION = 313

c_int_p = C.POINTER(C.c_int)
c_i64_p = C.POINTER(C.c_int64)
c_size_p = C.POINTER(C.c_size_t)
vp = C.c_void_p

# name -> (restype, argtypes); mirrors include/ompl_amd.h
SIGNATURES = {
    "ompl_amd_init": (C.c_int, [C.c_int]),
    "ompl_amd_last_error": (C.c_char_p, []),
    "ompl_amd_version": (C.c_int, []),
    "ompl_amd_synchronize": (C.c_int, [vp]),
    "ompl_amd_capabilities": (C.c_int, [vp]),
    "ompl_amd_malloc": (C.c_int, [C.POINTER(vp), C.c_size_t]),
    "ompl_amd_free": (C.c_int, [vp]),
    "ompl_amd_memcpy_h2d": (C.c_int, [vp, vp, C.c_size_t]),
    "ompl_amd_memcpy_d2h": (C.c_int, [vp, vp, C.c_size_t]),
    "ompl_amd_env_create": (C.c_int, [vp, C.c_int, vp, C.c_int, C.POINTER(vp)]),
    "ompl_amd_env_destroy": (C.c_int, [vp]),
    "ompl_amd_env_bytes": (C.c_int, [vp, c_size_p]),
    "ompl_amd_validate_states_dev": (C.c_int, [vp, vp, C.c_int64, vp, vp]),
    "ompl_amd_validate_states_f64_dev": (C.c_int, [vp, vp, C.c_int64, vp, vp]),
    "ompl_amd_is_valid": (C.c_int, [vp, vp, c_int_p]),
    "ompl_amd_arm_spheres_dev": (C.c_int, [vp, C.c_int64, vp, vp]),
    "ompl_amd_check_motions_dev": (C.c_int, [vp, vp, vp, C.c_int64, C.c_double, C.c_uint, vp, vp, vp]),
    "ompl_amd_check_motions_noalloc_dev": (C.c_int, [vp, vp, vp, C.c_int64, C.c_double, C.c_uint, vp, vp, vp]),
    "ompl_amd_check_motions_last_valid_dev": (C.c_int, [vp, vp, vp, C.c_int64, C.c_double, C.c_uint, vp, vp,
                                                        vp]),
    "ompl_amd_check_motion": (C.c_int, [vp, vp, vp, C.c_double, C.c_uint, c_int_p]),
    "ompl_amd_check_motion_last_valid": (C.c_int, [vp, vp, vp, C.c_double, C.c_uint, c_int_p,
                                                   C.POINTER(C.c_double)]),
    "ompl_amd_check_state_lists_dev": (C.c_int, [vp, vp, C.c_int64, vp, C.c_int64, vp, vp, vp]),
    "ompl_amd_check_states": (C.c_int, [vp, vp, C.c_int, c_int_p, c_int_p]),
    "ompl_amd_motion_graph_create": (C.c_int, [vp, vp, vp, C.c_int64, C.c_double, C.c_uint, vp, vp, C.POINTER(vp)]),
    "ompl_amd_motion_graph_create_group": (C.c_int, [vp, C.c_int, vp, vp, vp, C.c_double, C.c_uint, vp, vp,
                                                     C.POINTER(vp)]),
    "ompl_amd_motion_graph_launch": (C.c_int, [vp, vp]),
    "ompl_amd_motion_graph_destroy": (C.c_int, [vp]),
    "ompl_amd_get_tuning": (C.c_int, [vp]),
    "ompl_amd_set_tuning": (C.c_int, [vp]),
    "ompl_amd_env_set_tuning": (C.c_int, [vp, vp]),
    "ompl_amd_env_get_tuning": (C.c_int, [vp, vp, C.POINTER(C.c_int)]),
    "ompl_amd_autotune_motions": (C.c_int, [vp, vp, vp, C.c_int64, C.c_double, C.c_uint, C.c_int, vp,
                                            C.POINTER(C.c_double)]),
    "ompl_amd_autotune_motions_streams": (C.c_int, [vp, vp, vp, C.c_int64, C.c_double, C.c_uint, vp, C.c_int,
                                                   C.c_int, vp, vp]),
    "ompl_amd_motion_work": (C.c_int, [vp, vp, vp, C.c_int64, C.c_double, C.c_uint, vp]),
    "ompl_amd_validity_work": (C.c_int, [vp, vp, C.c_int64, vp]),
    "ompl_amd_profile_enable": (C.c_int, [C.c_int]),
    "ompl_amd_profile_reset": (C.c_int, []),
    "ompl_amd_profile_query": (C.c_int, [C.c_char_p, c_i64_p, C.POINTER(C.c_double)]),
    "ompl_amd_sample_uniform_dev": (C.c_int, [vp, C.c_int, C.c_int64, C.c_int, vp, vp, vp, vp, vp]),
    "ompl_amd_sample_valid_dev": (C.c_int, [vp, vp, C.c_int, C.c_int64, C.c_int, vp, vp, vp, vp, vp, vp]),
    "ompl_amd_rng_create": (C.c_int, [vp, C.c_int, C.POINTER(vp)]),
    "ompl_amd_rng_destroy": (C.c_int, [vp]),
    "ompl_amd_rng_sample_uniform_dev": (C.c_int, [vp, C.c_int64, C.c_int, vp, vp, vp, vp, vp]),
    "ompl_amd_rng_sample_se3_dev": (C.c_int, [vp, C.c_int64, vp, vp, vp, vp, vp]),
    "ompl_amd_rng_sample_valid_dev": (C.c_int, [vp, vp, C.c_int64, C.c_int, vp, vp, vp, vp, vp, vp]),
    "ompl_amd_nn_create": (C.c_int, [C.c_int, C.c_int, C.POINTER(vp)]),
    "ompl_amd_nn_destroy": (C.c_int, [vp]),
    "ompl_amd_nn_clear": (C.c_int, [vp]),
    "ompl_amd_nn_add": (C.c_int, [vp, vp, C.c_int64, c_i64_p]),
    "ompl_amd_nn_add_dev": (C.c_int, [vp, vp, C.c_int64, c_i64_p, vp]),
    "ompl_amd_nn_remove": (C.c_int, [vp, C.c_int64, c_int_p]),
    "ompl_amd_nn_size": (C.c_int, [vp, c_i64_p]),
    "ompl_amd_nn_build": (C.c_int, [vp, vp]),
    "ompl_amd_nn_knn_dev": (C.c_int, [vp, vp, C.c_int64, C.c_int, vp, vp, vp]),
    "ompl_amd_nn_radius_dev": (C.c_int, [vp, vp, C.c_int64, C.c_float, vp, vp, vp, C.c_int64, c_i64_p, vp]),
    "ompl_amd_nn_knn_packed_dev": (C.c_int, [vp, vp, C.c_int64, vp, C.c_int, vp, vp, vp, vp]),
    "ompl_amd_nn_knn_packed": (C.c_int, [vp, vp, C.c_int64, vp, C.c_int, vp, vp, vp]),
    "ompl_amd_nn_radius_packed": (C.c_int, [vp, vp, C.c_int64, vp, vp, vp, vp, C.c_int64, c_i64_p]),
    "ompl_amd_nn_radius_packed_dev": (C.c_int, [vp, vp, C.c_int64, vp, vp, vp, vp, C.c_int64, c_i64_p, vp]),
    "ompl_amd_nn_knn": (C.c_int, [vp, vp, C.c_int64, C.c_int, vp, vp]),
    "ompl_amd_nn_radius": (C.c_int, [vp, vp, C.c_int64, C.c_float, vp, vp, vp, C.c_int64, c_i64_p]),
    "ompl_amd_nn_nearest": (C.c_int, [vp, vp, c_i64_p]),
    "ompl_amd_rrtstar_radius": (C.c_double, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int64]),
    "ompl_amd_server_create": (C.c_int, [vp, C.POINTER(vp)]),
    "ompl_amd_server_destroy": (C.c_int, [vp]),
    "ompl_amd_server_submit": (C.c_int, [vp, vp, c_i64_p]),
    "ompl_amd_server_poll": (C.c_int, [vp, C.c_int64, vp]),
    "ompl_amd_server_wait": (C.c_int, [vp, C.c_int64, C.c_double, vp]),
    "ompl_amd_server_path": (C.c_int, [vp, C.c_int64, vp, C.c_int, c_int_p]),
    "ompl_amd_server_cancel": (C.c_int, [vp, C.c_int64]),
    "ompl_amd_server_release": (C.c_int, [vp, C.c_int64]),
    "ompl_amd_server_stats": (C.c_int, [vp, vp]),
    "ompl_amd_server_footprint": (C.c_size_t, [C.c_int]),
    "ompl_amd_server_phase_us": (C.c_int, [vp, vp]),
    "ompl_amd_plan_plain": (C.c_int, [vp, vp, vp, C.c_int, c_int_p]),
    "ompl_amd_prm_build": (C.c_int, [vp, C.c_uint32, C.c_int, C.c_int64, C.c_int, C.c_double, C.c_uint, C.c_uint,
                                     C.POINTER(vp), vp]),
    "ompl_amd_prm_destroy": (C.c_int, [vp]),
    "ompl_amd_prm_milestones": (C.c_int, [vp, vp, vp, C.c_int64]),
    "ompl_amd_prm_edges": (C.c_int, [vp, vp, vp, vp, C.c_int64]),
    "ompl_amd_prm_components": (C.c_int, [vp, vp, C.c_int64]),
    "ompl_amd_prm_graph": (C.c_int, [vp, vp, vp, C.c_int64]),
    "ompl_amd_workload_obstacles": (C.c_int, [C.c_uint32, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                              C.c_double, C.c_double, vp, vp]),
    "ompl_amd_workload_configs": (C.c_int, [C.c_uint32, C.c_int64, vp]),
    "ompl_amd_workload_edges": (C.c_int, [C.c_uint32, C.c_int64, C.c_double, C.c_double, C.c_double,
                                          C.c_double, vp, vp]),
    "ompl_amd_workload_se3": (C.c_int, [C.c_uint32, C.c_int64, C.c_double, C.c_double, vp]),
}


class OmplAmdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Tuning(C.Structure):
    _fields_ = [
        ("validity_lanes_per_state", C.c_int),
        ("validity_obstacle_parts", C.c_int),
        ("validity_block", C.c_int),
        ("motion_lanes_per_state", C.c_int),
        ("motion_obstacle_parts", C.c_int),
        ("motion_block", C.c_int),
        ("motion_waves_per_cu", C.c_int),
        ("motion_strategy", C.c_int),
        ("validity_shared_fk", C.c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Caps(C.Structure):
    _fields_ = [
        ("arm_dof", C.c_int),
        ("arm_spheres", C.c_int),
        ("max_env_words", C.c_int),
        ("max_segments", C.c_int),
        ("max_k", C.c_int),
        ("max_nn_dim", C.c_int),
        ("max_sample_dim", C.c_int),
        ("device", C.c_int),
        ("compute_units", C.c_int),
        ("device_name", C.c_char * 64),
    ]


class PrmStats(C.Structure):
    _fields_ = [
        ("samples", C.c_int64),
        ("milestones", C.c_int64),
        ("candidate_edges", C.c_int64),
        ("valid_edges", C.c_int64),
        ("components", C.c_int64),
        ("k", C.c_int64),
        ("ms_sample", C.c_double),
        ("ms_validity", C.c_double),
        ("ms_knn", C.c_double),
        ("ms_candidates", C.c_double),
        ("ms_motions", C.c_double),
        ("ms_graph", C.c_double),
        ("ms_total", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def load(path: str | os.PathLike | None = None):
    """Loads libompl_amd.so (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    # torch ships its own HIP runtime under the same soname as ROCm's: load it
    # first (when present) so that this library and torch share one runtime,
    # whichever module the program imports first
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not p.exists():
        raise OmplAmdError(-8, f"{p} is not built; run `python -m ompl_amd.build` (hipcc, gfx950)")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    got = lib.ompl_amd_version()
    if got != ABI_VERSION:
        raise OmplAmdError(-8, f"{p} has ABI version {got}, these bindings need {ABI_VERSION}: rebuild the library "
                               "(python -m ompl_amd.build)")
    if path is None:
        _lib = lib
    return lib


def exported_symbols(lib=None) -> dict[str, bool]:
    lib = lib or load()
    return {name: hasattr(lib, name) for name in SIGNATURES}


def check(code: int) -> None:
    if code != 0:
        msg = load().ompl_amd_last_error()
        raise OmplAmdError(code, msg.decode() if msg else "unknown error")


def call(name: str, *args):
    fn = getattr(load(), name)
    check(fn(*args))
