"""CPU pins of the oracle's batched PRM restatement: component labels equal
scipy's connected components (each labelled by its smallest member), the
candidate set equals a numpy brute-force kNN union computed in FP64 where no
near-ties exist, samples replay std::mt19937, and edge validity equals the
per-edge motion check."""
import numpy as np


def test_disjoint_set_labels_match_scipy(oracle):
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components

    rng = np.random.default_rng(3)
    for n, m in [(1, 0), (5, 0), (50, 30), (400, 380), (2000, 1500)]:
        e = rng.integers(0, n, size=(m, 2)).astype(np.int32)
        valid = rng.random(m) < 0.7
        lab = oracle.disjoint_set_labels(n, e, valid)
        ev = e[valid]
        g = coo_matrix((np.ones(len(ev)), (ev[:, 0], ev[:, 1])), shape=(n, n))
        nc, sl = connected_components(g, directed=False)
        first = np.full(nc, n, np.int64)
        np.minimum.at(first, sl, np.arange(n))
        assert np.array_equal(lab, first[sl])


def test_prm_batch_structure(oracle):
    from ompl_amd import workload  # host-side generator, no device needed

    s, b = workload.obstacles(1007, 16, 16)
    env = oracle.OracleEnv(s, b)
    k = 6
    r = oracle.prm_batch(env, 7, 600, k, 0.01, 1, 3, threads=4)
    # samples: stream s replays mt19937(7 + s)
    lo, hi = np.full(7, -np.pi), np.full(7, np.pi)
    assert np.array_equal(r["samples"][200:400], oracle.Mt19937(8).sample_uniform(lo, hi, 200))
    # milestones: valid samples in order
    for i in (0, 1, 17, 599):
        assert r["valid"][i] == oracle.state_valid(env, r["samples"][i].astype(np.float32))
    nodes = r["milestones"]
    nv = nodes.shape[0]
    assert 50 < nv < 600
    # candidates: undirected union of k nearest other milestones (FP64 brute force)
    d = ((nodes[:, None, :].astype(np.float32).astype(np.float64)
          - nodes[None, :, :].astype(np.float32).astype(np.float64)) ** 2).sum(-1)
    np.fill_diagonal(d, np.inf)
    order = np.argsort(d, axis=1, kind="stable")
    keys = set()
    for i in range(nv):
        gap = d[i, order[i, k]] - d[i, order[i, k - 1]]
        assert gap > 1e-5 * d[i, order[i, k]]  # no near-tie at the k-th neighbour for this seed
        keys.update((min(i, j), max(i, j)) for j in order[i, :k].tolist())
    assert keys == set(map(tuple, r["edges"].tolist()))
    # edge validity: per-edge motion check from the lower index
    for t in range(0, r["edges"].shape[0], 97):
        u, v = r["edges"][t]
        ok, _ = oracle.check_motion(env, nodes[u], nodes[v], 0.01)
        assert ok == r["edge_valid"][t]


def test_kstar_k_known_values(oracle):
    # k = ceil((e + e/7) ln n): e(1 + 1/7) = 3.1066078..., ln 1000 = 6.9077553 -> 21.46 -> 22;
    # ln 135102 = 11.813786 -> 36.70 -> 37; ln 2 = 0.693 -> 2.15 -> 3; n = 1 -> 0, raised to 1
    assert oracle.kstar_k(1000) == 22
    assert oracle.kstar_k(135102) == 37
    assert oracle.kstar_k(2) == 3
    assert oracle.kstar_k(1) == 1
    assert oracle.kstar_k(10 ** 12) == 63


def test_roadmap_shortest_path_against_scipy(oracle):
    """The oracle's shortest path on a roadmap (PRM request answers) is a
    chain of valid edges from milestone 0 to 1 whose FP64 length equals
    scipy's Dijkstra distance; unreachable goal -> empty."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import dijkstra

    rng = np.random.default_rng(4)
    for trial in range(6):
        nv = 300
        nodes = rng.uniform(-3, 3, (nv, 7))
        e = set()
        for i in range(nv):
            for j in rng.choice(nv, 4, replace=False):
                if i != j:
                    e.add((min(i, int(j)), max(i, int(j))))
        edges = np.array(sorted(e), np.int32)
        valid = rng.uniform(size=len(edges)) < (0.8 if trial < 5 else 0.05)
        w = np.sqrt(((nodes[edges[:, 0]] - nodes[edges[:, 1]]) ** 2).sum(1))
        ev = edges[valid]
        g = coo_matrix((w[valid], (ev[:, 0], ev[:, 1])), shape=(nv, nv))
        dist = dijkstra(g, directed=False, indices=0)
        path = oracle.roadmap_shortest_path(nv, edges, valid, nodes)
        if not np.isfinite(dist[1]):
            assert path == []
            continue
        assert path[0] == 0 and path[-1] == 1
        keys = set(map(tuple, ev.tolist()))
        length = 0.0
        for a, b in zip(path[:-1], path[1:]):
            assert (min(a, b), max(a, b)) in keys
            length += float(np.sqrt(((nodes[a] - nodes[b]) ** 2).sum()))
        assert abs(length - dist[1]) < 1e-9 * max(1.0, dist[1])
