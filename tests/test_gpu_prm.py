"""GPU parity of the batched PRM roadmap (samples, milestones, candidate
edges, motion validity, components, CSR graph) against the CPU oracle's
restatement of PRM::growRoadmap / addMilestone with KStrategy: bit-identical
samples, identical index sets and labels."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def csr_of(nv, edges, valid):
    """Symmetric CSR of the valid edges, neighbours ascending."""
    e = edges[valid]
    rows = np.concatenate([e[:, 0], e[:, 1]])
    cols = np.concatenate([e[:, 1], e[:, 0]])
    order = np.lexsort((cols, rows))
    off = np.zeros(nv + 1, np.int64)
    np.add.at(off, rows + 1, 1)
    return np.cumsum(off), cols[order].astype(np.int32)


def roadmap_pair(gpu, oracle, env_seed, ns, nb, seed, n, k, streams, lvs=0.01):
    from ompl_amd.prm import build_roadmap

    s, b = gpu.workload.obstacles(env_seed, ns, nb)
    env = gpu.Env(s, b)
    print(f"prm env {env_seed} ({ns}+{nb}) seed {seed} n {n} k {k} streams {streams}", flush=True)
    g = build_roadmap(env, seed, n, k=k, lvs=lvs, n_streams=streams)
    o = oracle.prm_batch(oracle.OracleEnv(s, b), seed, n, k, lvs, 1, streams, threads=8)
    return g, o


def assert_same(g, o):
    assert g.stats["samples"] == o["samples"].shape[0]
    assert np.array_equal(g.sample_index, o["sample_index"])
    assert np.array_equal(g.milestones, o["milestones"])
    assert np.array_equal(g.edges, o["edges"])
    assert np.array_equal(g.edge_valid, o["edge_valid"])
    assert np.array_equal(g.component, o["component"])
    off, adj = csr_of(g.milestones.shape[0], o["edges"], o["edge_valid"])
    assert np.array_equal(g.offsets, off) and np.array_equal(g.adjacency, adj)
    st = g.stats
    assert st["milestones"] == g.milestones.shape[0] and st["candidate_edges"] == g.edges.shape[0]
    assert st["valid_edges"] == int(g.edge_valid.sum())
    assert st["components"] == len(set(g.component.tolist()))


@pytest.mark.parametrize("streams", [1, 7, 64])
def test_prm_matches_oracle(gpu, oracle, streams):
    g, o = roadmap_pair(gpu, oracle, 1007, 16, 16, 7, 3000, 10, streams)
    assert 0 < g.edge_valid.sum() < g.edges.shape[0]  # both outcomes present
    assert_same(g, o)


@pytest.mark.parametrize("k", [1, 3, 63])
def test_prm_k_values(gpu, oracle, k):
    g, o = roadmap_pair(gpu, oracle, 1011, 8, 8, 70 + k, 700, k, 5)
    assert_same(g, o)


def test_prm_fewer_milestones_than_k(gpu, oracle):
    # 40 samples among 32 obstacles: a handful of milestones, k larger than the set
    g, o = roadmap_pair(gpu, oracle, 1007, 16, 16, 9, 40, 30, 2)
    assert g.milestones.shape[0] < 31
    assert_same(g, o)


def test_prm_no_milestones(gpu, oracle):
    # 1,024 obstacles: no sampled configuration is collision free
    g, o = roadmap_pair(gpu, oracle, 1004, 512, 512, 7, 256, 10, 4)
    assert g.milestones.shape[0] == o["milestones"].shape[0]
    assert_same(g, o)


def test_prm_components_against_scipy(gpu):
    """At a size the oracle does not replay: labels are the connected components
    of the valid-edge graph (scipy), each labelled by its smallest member; CSR
    symmetric and sorted; candidate edges are exactly the undirected kNN union."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components

    from ompl_amd.nn import NearestNeighborsGPU
    from ompl_amd.prm import build_roadmap

    s, b = gpu.workload.obstacles(1001, 32, 32)
    env = gpu.Env(s, b)
    g = build_roadmap(env, 7, 60000, k=10, n_streams=256)
    nv = g.milestones.shape[0]
    assert nv > 10000
    e = g.edges[g.edge_valid]
    m = coo_matrix((np.ones(len(e)), (e[:, 0], e[:, 1])), shape=(nv, nv))
    nc, lab = connected_components(m, directed=False)
    assert nc == g.stats["components"]
    first = np.full(nc, nv, np.int64)
    np.minimum.at(first, lab, np.arange(nv))
    assert np.array_equal(g.component, first[lab])
    off, adj = csr_of(nv, g.edges, g.edge_valid)
    assert np.array_equal(g.offsets, off) and np.array_equal(g.adjacency, adj)
    nn = NearestNeighborsGPU(0, 7)
    f32 = g.milestones.astype(np.float32)
    nn.add(f32)
    idx, _ = nn.knn_batch(f32, 11)
    keys = set()
    for i in range(nv):
        nb = [j for j in idx[i].tolist() if j != i][:10]
        keys.update((min(i, j), max(i, j)) for j in nb)
    assert keys == set(map(tuple, g.edges.tolist()))


def test_prm_autotune_keeps_winner_on_environment(gpu, oracle):
    """The autotune flag tunes the motion checks on the roadmap's candidate
    edges and keeps the winner as the environment's own shapes: the library's
    tuning is unchanged, and the roadmap still equals the oracle's."""
    from ompl_amd.prm import build_roadmap

    s, b = gpu.workload.obstacles(1000, 24, 24)
    env = gpu.Env(s, b)
    before = gpu.get_tuning()
    g = build_roadmap(env, 5, 3000, k=10, lvs=0.01, n_streams=16, autotune=True)
    assert gpu.get_tuning() == before
    assert env.tuning()["own"] is True
    o = oracle.prm_batch(oracle.OracleEnv(s, b), 5, 3000, 10, 0.01, 1, 16, threads=8)
    assert_same(g, o)
    env.close()


def test_prm_argument_errors(gpu):
    from ompl_amd._lib import OmplAmdError
    from ompl_amd.prm import build_roadmap

    s, b = gpu.workload.obstacles(1007, 4, 4)
    env = gpu.Env(s, b)
    for kw in (dict(k=0), dict(k=64), dict(lvs=0.0), dict(n_streams=0)):
        with pytest.raises(OmplAmdError):
            build_roadmap(env, 1, 100, **kw)


def test_prm_kstar_matches_oracle(gpu, oracle):
    from ompl_amd.prm import build_roadmap

    s, b = gpu.workload.obstacles(1007, 16, 16)
    env = gpu.Env(s, b)
    g = build_roadmap(env, 11, 2500, n_streams=16, kstar=True)
    o = oracle.prm_batch(oracle.OracleEnv(s, b), 11, 2500, 10, 0.01, 1, 16, threads=8, kstar=True)
    assert g.stats["k"] == o["k"] == oracle.kstar_k(g.milestones.shape[0])
    assert_same(g, o)


def test_prm_config4_literal_scene_full_size(gpu, oracle):
    """Config 4 as stated: 200,192 samples (256 generators) with 512 spheres +
    512 boxes of the motion benchmark's sizes.  That scene is degenerate: the
    obstacles fill the arm's reach, so no sample is collision free; the full
    sample set's validity is checked against the oracle and the roadmap is
    empty."""
    g, o = roadmap_pair(gpu, oracle, 1003, 512, 512, 7, 200_192, 10, 256)
    assert o["samples"].shape[0] == 200_192 and not o["valid"].any()
    assert g.milestones.shape[0] == 0
    assert_same(g, o)


def test_prm_config4_small_obstacles_full_size(gpu, oracle):
    """200,192 samples among 1,024 small obstacles (radii and half-extents
    U[0.01, 0.03]; about 9 % of samples are collision free): milestones equal
    the oracle's over the whole sample set; candidate edges are the undirected
    kNN union (checked for a spread of milestones and of edges with the
    oracle's kNN); edge validity equals the oracle's on a spread of 4,096
    edges; components equal scipy's on the roadmap's own edges."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components

    from ompl_amd.prm import build_roadmap

    s, b = gpu.workload.obstacles(1003, 512, 512, r=(0.01, 0.03), h=(0.01, 0.03))
    env = gpu.Env(s, b)
    oenv = oracle.OracleEnv(s, b)
    print("prm config 4, small obstacles, 200192 samples", flush=True)
    g = build_roadmap(env, 7, 200_192, k=10, n_streams=256)
    per = 200_192 // 256
    lo, hi = np.full(7, -np.pi), np.full(7, np.pi)
    samples = np.concatenate([oracle.Mt19937(7 + s_).sample_uniform(lo, hi, per) for s_ in range(256)])
    valid = oracle.validate_states_mt(oenv, samples.T.astype(np.float32), 16).astype(bool)
    sidx = np.nonzero(valid)[0].astype(np.int32)
    assert np.array_equal(g.sample_index, sidx) and np.array_equal(g.milestones, samples[sidx])
    nv = len(sidx)
    assert nv > 10_000
    edges = g.edges
    keyset = set(map(tuple, edges.tolist()))
    f32 = g.milestones.astype(np.float32)
    rows = np.linspace(0, nv - 1, 512).astype(int)
    oi, _ = oracle.knn_mt(0, f32, f32[rows], 11, 16)
    for i, nbr in zip(rows, oi):
        for j in [j for j in nbr.tolist() if j != i][:10]:
            assert (min(i, j), max(i, j)) in keyset
    pick = np.linspace(0, len(edges) - 1, 2048).astype(int)
    ends = np.unique(edges[pick].ravel())
    ei, _ = oracle.knn_mt(0, f32, f32[ends], 11, 16)
    nbrs = {int(u): set(j for j in row.tolist() if j != u) for u, row in zip(ends, ei)}
    for u, v in edges[pick]:
        assert v in nbrs[int(u)] or u in nbrs[int(v)]
    sel = np.linspace(0, len(edges) - 1, 4096).astype(int)
    ov, _ = oracle.check_motions(oenv, g.milestones[edges[sel, 0]], g.milestones[edges[sel, 1]], 0.01, 1, threads=16)
    assert np.array_equal(g.edge_valid[sel], ov.astype(bool))
    assert 0 < g.edge_valid.sum() < len(edges)
    e = edges[g.edge_valid]
    nc, lab = connected_components(coo_matrix((np.ones(len(e)), (e[:, 0], e[:, 1])), shape=(nv, nv)), directed=False)
    first = np.full(nc, nv, np.int64)
    np.minimum.at(first, lab, np.arange(nv))
    assert nc == g.stats["components"] and np.array_equal(g.component, first[lab])
