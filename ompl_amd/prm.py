"""Batched PRM roadmap over the device pipeline of libompl_amd.

Mirrors the inner loop of ompl::geometric::PRM (geometric/planners/prm/src/
PRM.cpp:364-385 growRoadmap, :570-604 addMilestone) with the KStrategy
connection rule (geometric/planners/prm/ConnectionStrategy.h:56-81) and the
default k = magic::DEFAULT_NEAREST_NEIGHBORS = 10 (PRM.cpp:68):

* milestones are the valid samples, in sample order;
* candidate edges are the undirected union of every milestone's k nearest
  milestones (exact FP32 L2, equal distances ordered by lower index);
* each candidate (u < v) is motion-checked from milestone u to milestone v
  (the reference checks from the older milestone n to the new one m,
  PRM.cpp:590);
* components follow PRM::uniteComponents (PRM.cpp:606-609); a component is
  labelled by its smallest milestone index.

Batch semantics: the reference's addMilestone queries only the milestones
added before m, so its edge set depends on insertion order; here every
milestone queries all milestones.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import PrmStats, call

__all__ = ["Roadmap", "build_roadmap", "DEFAULT_NEAREST_NEIGHBORS"]

DEFAULT_NEAREST_NEIGHBORS = 10


@dataclass
class Roadmap:
    stats: dict
    milestones: np.ndarray      # [nv, 7] float64
    sample_index: np.ndarray    # [nv] int32
    edges: np.ndarray           # [ne, 2] int32, u < v, sorted
    edge_valid: np.ndarray      # [ne] bool
    component: np.ndarray       # [nv] int32 (smallest milestone index of the component)
    offsets: np.ndarray         # [nv + 1] int64 CSR over the valid edges
    adjacency: np.ndarray       # [2 * valid_edges] int32

    def neighbors(self, i: int) -> np.ndarray:
        return self.adjacency[self.offsets[i]:self.offsets[i + 1]]

    def same_component(self, a: int, b: int) -> bool:
        """PRM::sameComponent (PRM.cpp:611-614): same label."""
        return bool(self.component[a] == self.component[b])


def build_roadmap(env, seed: int, n_samples: int, k: int = DEFAULT_NEAREST_NEIGHBORS, lvs: float = 0.01,
                  factor: int = 1, n_streams: int = 64, fetch: bool = True,
                  autotune: bool = False, kstar: bool = False) -> Roadmap | dict:
    """Builds the roadmap on the device.  lvs is the longest valid segment in
    radians (the motion resolution); samples come from n_streams generators
    seeded seed, seed + 1, ...  autotune runs the motion autotuner on the
    candidate edges first (the winner becomes the environment's own launch
    shapes; the library's tuning is left as it was).  kstar uses PRM*'s
    KStarStrategy, k = ceil((e + e/7) ln(milestones)) (ConnectionStrategy.h:
    125-157), instead of k.  With fetch=False only the statistics are
    returned."""
    h = C.c_void_p()
    st = PrmStats()
    call("ompl_amd_prm_build", env.handle, int(seed) & 0xffffffff, int(n_streams), int(n_samples), int(k),
         float(lvs), int(factor), (1 if autotune else 0) | (2 if kstar else 0), C.byref(h), C.byref(st))
    try:
        stats = st.as_dict()
        if not fetch:
            return stats
        nv, ne, nvalid = st.milestones, st.candidate_edges, st.valid_edges
        nodes = np.zeros((nv, 7), np.float64)
        sidx = np.zeros(nv, np.int32)
        call("ompl_amd_prm_milestones", h, nodes.ctypes.data_as(C.c_void_p), sidx.ctypes.data_as(C.c_void_p), nv)
        u = np.zeros(ne, np.int32)
        v = np.zeros(ne, np.int32)
        ok = np.zeros(ne, np.uint8)
        call("ompl_amd_prm_edges", h, u.ctypes.data_as(C.c_void_p), v.ctypes.data_as(C.c_void_p),
             ok.ctypes.data_as(C.c_void_p), ne)
        lab = np.zeros(nv, np.int32)
        if nv:
            call("ompl_amd_prm_components", h, lab.ctypes.data_as(C.c_void_p), nv)
        off = np.zeros(nv + 1, np.int64)
        adj = np.zeros(2 * nvalid, np.int32)
        call("ompl_amd_prm_graph", h, off.ctypes.data_as(C.c_void_p), adj.ctypes.data_as(C.c_void_p), adj.shape[0])
        return Roadmap(stats, nodes, sidx, np.stack([u, v], 1), ok.astype(bool), lab, off, adj)
    finally:
        call("ompl_amd_prm_destroy", h)
