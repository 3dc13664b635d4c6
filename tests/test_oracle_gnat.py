"""Pins the oracle's NearestNeighborsGNAT restatement (oracle/gnat.cpp) on the
CPU, the way the reference tests GNAT (tests/datastructures/nearestneighbors.cpp:
equal results to a linear scan):
  * against the reference's own NearestNeighborsLinear, compiled unchanged
    from the reference tree (tests/golden/nn_linear_*.npz): k-NN distance
    lists equal bit for bit (both use FP64 distances of the same FP32
    coordinates), index lists equal up to the order of equal distances
    (which partial_sort leaves undefined), radius lists equal;
  * against an FP64 brute force at sizes large enough to split the tree
    several levels deep (R^7 and SE3, many seeds for the pivot draws),
    including duplicate points and k larger than the set."""
import math
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"


def brute(metric, data, q):
    """FP64 distances with the terms summed in index order (as the C code)."""
    x = data.astype(np.float64)[None, :, :]
    y = q.astype(np.float64)[:, None, :]

    def seq_sum(terms):
        acc = np.zeros(terms[0].shape)
        for t in terms:
            acc = acc + t
        return acc

    if metric == 0:
        return np.sqrt(seq_sum([(x[..., k] - y[..., k]) ** 2 for k in range(x.shape[-1])]))
    t = np.sqrt(seq_sum([(x[..., k] - y[..., k]) ** 2 for k in range(3)]))
    dot = np.abs(seq_sum([x[..., k] * y[..., k] for k in range(3, 7)]))
    # libm's acos (math.acos), as the C code calls; numpy's vectorised arccos
    # may differ from it in the last bit
    rot = np.where(dot > 1.0 - 1e-9, 0.0, np.vectorize(math.acos)(np.minimum(dot, 1.0)))
    return t + rot


def same_up_to_ties(gi, gd, ri, rd):
    """Equal distances, and equal index sets within each run of equal distances."""
    if not np.array_equal(gd, rd):
        return False
    for dv in np.unique(rd):
        if set(gi[gd == dv].tolist()) != set(ri[rd == dv].tolist()):
            return False
    return True


@pytest.mark.parametrize("name,metric", [("nn_linear_l2", 0), ("nn_linear_se3", 1)])
def test_gnat_equals_reference_linear(oracle, name, metric):
    f = np.load(GOLDEN / f"{name}.npz")
    data, queries, k, r = f["data"], f["queries"], int(f["k"]), float(f["radius"])
    for seed in (1, 7, 12345):
        g = oracle.Gnat(metric, data, seed)
        gi, gd = g.knn(queries, k)
        for j in range(len(queries)):
            # the reference's last rank may sit in a tie with points beyond k
            kth = f["knn_dist"][j][-1]
            inner = f["knn_dist"][j] < kth
            assert np.array_equal(gd[j], f["knn_dist"][j]), (seed, j)
            assert same_up_to_ties(gi[j][inner], gd[j][inner], f["knn_idx"][j][inner], f["knn_dist"][j][inner])
        off, idx = g.radius(queries, r)
        for j in range(len(queries)):
            want = f["rad_idx"][f["rad_off"][j]:f["rad_off"][j + 1]]
            assert sorted(idx[off[j]:off[j + 1]].tolist()) == sorted(want.tolist()), (seed, j)
        assert np.array_equal(f["rad_off"], off)


@pytest.mark.parametrize("metric,n,dim", [(0, 20000, 7), (1, 20000, 7), (0, 5000, 3)])
def test_gnat_equals_brute_force(oracle, metric, n, dim):
    rng = np.random.default_rng(5 + metric)
    if metric == 0:
        data = rng.uniform(-np.pi, np.pi, (n, dim)).astype(np.float32)
        data[-50:] = data[:50]  # exact duplicates
        q = rng.uniform(-np.pi, np.pi, (40, dim)).astype(np.float32)
        r = 1.8 if dim == 7 else 0.6
    else:
        t = rng.uniform(-10, 10, (n, 3))
        qq = rng.normal(size=(n, 4))
        qq /= np.linalg.norm(qq, axis=1, keepdims=True)
        data = np.concatenate([t, qq], 1).astype(np.float32)
        q = data[rng.integers(0, n, 40)] + rng.normal(scale=0.05, size=(40, 7)).astype(np.float32)
        q[:, 3:] /= np.linalg.norm(q[:, 3:], axis=1, keepdims=True)
        q = q.astype(np.float32)
        r = 2.5
    d = brute(metric, data, q)
    for seed in (3, 99):
        g = oracle.Gnat(metric, data, seed)
        gi, gd = g.knn(q, 16, threads=4)
        for j in range(len(q)):
            order = np.lexsort((np.arange(n), d[j]))[:16]
            assert np.array_equal(gd[j], d[j][order]), (seed, j)
            assert np.array_equal(gi[j], order), (seed, j)
        off, idx = g.radius(q, r, threads=4)
        for j in range(len(q)):
            want = np.nonzero(d[j] <= r)[0]
            want = want[np.lexsort((want, d[j][want]))]
            assert np.array_equal(idx[off[j]:off[j + 1]], want), (seed, j)


def test_gnat_k_larger_than_set(oracle):
    data = np.random.default_rng(1).uniform(-1, 1, (30, 7)).astype(np.float32)
    g = oracle.Gnat(0, data, 1)
    gi, gd = g.knn(data[:3], 40)
    assert np.all(gi[:, :30] >= 0) and np.all(gi[:, 30:] == -1) and np.all(np.isinf(gd[:, 30:]))
    assert np.array_equal(np.sort(gi[0, :30]), np.arange(30))
