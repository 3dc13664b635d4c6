"""GPU parity of the nearest-neighbour index against the CPU oracle (exact:
identical index lists and bit-identical FP32 distances, ties by lower index)
and against the reference's NearestNeighborsLinear golden data (index sets
equal except at FP32 near-ties)."""
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def make_nn(gpu, metric, dim, data, batches=1):
    from ompl_amd.nn import NearestNeighborsGPU

    nn = NearestNeighborsGPU(metric, dim)
    for part in np.array_split(data, batches):
        nn.add(part)
    return nn


def se3_points(gpu, seed, n, lo=-10.0, hi=10.0):
    return gpu.workload.se3(seed, n, lo, hi)


@pytest.mark.parametrize("name,metric,dim", [("nn_linear_l2.npz", 0, 7), ("nn_linear_se3.npz", 1, 7)])
def test_knn_golden_reference_linear(gpu, oracle, name, metric, dim):
    f = np.load(GOLDEN / name)
    data, queries, k = f["data"], f["queries"], int(f["k"])
    nn = make_nn(gpu, metric, dim, data)
    gi, gd = nn.knn_batch(queries, k)
    oi, od = oracle.knn(metric, data, queries, k)
    assert np.array_equal(gi, oi) and np.array_equal(gd, od)
    ties = 0
    for j in range(queries.shape[0]):
        if set(gi[j].tolist()) != set(f["knn_idx"][j].tolist()):
            ties += 1  # only allowed at an FP32 near-tie of the k-th distance
            assert abs(float(gd[j][-1]) - float(f["knn_dist"][j][-1])) < 2e-5
    assert ties <= 2


@pytest.mark.parametrize("name,metric,dim", [("nn_linear_l2.npz", 0, 7), ("nn_linear_se3.npz", 1, 7)])
def test_radius_golden_reference_linear(gpu, oracle, name, metric, dim):
    f = np.load(GOLDEN / name)
    data, queries, r = f["data"], f["queries"], float(f["radius"])
    nn = make_nn(gpu, metric, dim, data, batches=3)
    off, gi, gd = nn.radius_batch(queries, r)
    ores = oracle.radius(metric, data, queries, r)
    for j, (oi, od) in enumerate(ores):
        assert np.array_equal(gi[off[j]:off[j + 1]], oi)
        assert np.array_equal(gd[off[j]:off[j + 1]], od)


def test_knn_config1_r7(gpu, oracle):
    """100k states U[-pi,pi]^7, k=10 (4096 queries on the GPU; 384 checked
    against the oracle's full scan)."""
    rng = np.random.default_rng(2)
    data = rng.uniform(-np.pi, np.pi, (100_000, 7)).astype(np.float32)
    queries = rng.uniform(-np.pi, np.pi, (4096, 7)).astype(np.float32)
    nn = make_nn(gpu, 0, 7, data, batches=2)
    gi, gd = nn.knn_batch(queries, 10)
    assert np.all(np.diff(gd, axis=1) >= 0)
    sel = np.arange(0, 4096, 4096 // 384)[:384]
    oi, od = oracle.knn(0, data, queries[sel], 10)
    assert np.array_equal(gi[sel], oi) and np.array_equal(gd[sel], od)


def test_knn_config3_se3_sample(gpu, oracle):
    """1M SE3 states, 65,536 queries, k=16; 48 queries checked exactly."""
    data = se3_points(gpu, 5, 1_000_000)
    queries = se3_points(gpu, 6, 65_536)
    nn = make_nn(gpu, 1, 7, data)
    gi, gd = nn.knn_batch(queries, 16)
    assert np.all(np.diff(gd, axis=1) >= 0) and np.all(gi >= 0)
    sel = np.linspace(0, 65_535, 48).astype(int)
    oi, od = oracle.knn(1, data, queries[sel], 16)
    assert np.array_equal(gi[sel], oi) and np.array_equal(gd[sel], od)


def test_radius_config3_se3_sample(gpu, oracle):
    """Config 3 at its stated shape: 1M SE3 states, 65,536 queries, radius from
    the RRT* schedule at n = 1e6; 48 queries spread over the batch checked
    exactly (index lists and FP32 distances) against the oracle's full scan."""
    data = se3_points(gpu, 5, 1_000_000)
    queries = se3_points(gpu, 6, 65_536)
    r = gpu.rrtstar_radius(6, 8000.0 * np.pi ** 2, 1_000_000)
    assert abs(r - 0.97753) < 1e-4  # RRT* schedule at n = 1e6
    nn = make_nn(gpu, 1, 7, data)
    off, gi, gd = nn.radius_batch(queries, r)
    assert off[-1] > 65_536 and np.all(np.diff(off) >= 0)
    sel = np.linspace(0, 65_535, 48).astype(int).tolist()
    ores = oracle.radius(1, data, queries[sel], r)
    for j, (oi, od) in zip(sel, ores):
        assert np.array_equal(gi[off[j]:off[j + 1]], oi)
        assert np.array_equal(gd[off[j]:off[j + 1]], od)


def test_ties_duplicates_and_k_edge_cases(gpu, oracle):
    rng = np.random.default_rng(9)
    base = rng.uniform(-1, 1, (50, 7)).astype(np.float32)
    data = np.concatenate([base, base, base[:10]])  # exact duplicates: ties by lower index
    nn = make_nn(gpu, 0, 7, data, batches=4)
    q = np.concatenate([base[:8], rng.uniform(-1, 1, (8, 7)).astype(np.float32)])
    for k in (1, 3, 16, 17, 64):
        gi, gd = nn.knn_batch(q, k)
        oi, od = oracle.knn(0, data, q, k)
        assert np.array_equal(gi, oi) and np.array_equal(gd, od), k
    # k larger than the structure: -1 / +inf padding
    small = make_nn(gpu, 0, 7, base[:5])
    gi, gd = small.knn_batch(q[:2], 8)
    assert np.all(gi[:, 5:] == -1) and np.all(np.isinf(gd[:, 5:]))
    assert set(gi[0, :5].tolist()) == set(range(5))


def test_knn_split_bounds_with_duplicates_and_removal(gpu, oracle):
    """64 queries over 20,480 points searched in many splits, which exchange
    bounds on each query's k-th key: with every point present ten times (ties
    broken by the lower id) and 500 ids removed, the lists equal the
    oracle's for k from 1 (one slot) to 64 (one slot per 2.5 lanes)."""
    rng = np.random.default_rng(77)
    base = rng.uniform(-1, 1, (2048, 7)).astype(np.float32)
    data = np.tile(base, (10, 1))
    nn = make_nn(gpu, 0, 7, data, batches=2)
    removed = rng.choice(len(data), 500, replace=False)
    for i in removed:
        nn.remove(int(i))
    keep = np.setdiff1d(np.arange(len(data)), removed)
    q = np.concatenate([base[:16], rng.uniform(-1, 1, (48, 7)).astype(np.float32)])
    for k in (1, 4, 10, 16, 40, 64):
        gi, gd = nn.knn_batch(q, k)
        oi, od = oracle.knn(0, data[keep], q, k)
        assert np.array_equal(gi, keep[oi]) and np.array_equal(gd, od), k


def test_empty_nearest_raises_and_removal(gpu, oracle):
    from ompl_amd._lib import OmplAmdError

    nn = make_nn(gpu, 0, 3, np.zeros((0, 3), np.float32))
    with pytest.raises(OmplAmdError, match="No elements found"):
        nn.nearest([0, 0, 0])
    gi, gd = nn.knn_batch(np.zeros((2, 3), np.float32), 4)
    assert np.all(gi == -1)
    pts = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0], [3, 0, 0]], np.float32)
    nn.add(pts)
    assert nn.nearest([0.9, 0, 0]) == 1
    assert nn.remove(1) and not nn.remove(1) and nn.size() == 3
    assert nn.nearest([0.9, 0, 0]) == 0  # equal distance to 0 and 2: lower id
    i, d = nn.nearestK([0.9, 0, 0], 3)
    assert list(i) == [0, 2, 3]
    i, d = nn.nearestR([1.0, 0, 0], 1.0)  # inclusive radius
    assert list(i) == [0, 2]
    nn.clear()
    assert nn.size() == 0


def test_incremental_adds_match_oracle(gpu, oracle):
    rng = np.random.default_rng(21)
    nn = make_nn(gpu, 0, 7, np.zeros((0, 7), np.float32))
    data = np.zeros((0, 7), np.float32)
    for step in range(6):
        new = rng.uniform(-np.pi, np.pi, (int(rng.integers(1, 700)), 7)).astype(np.float32)
        nn.add(new)
        data = np.concatenate([data, new])
        q = rng.uniform(-np.pi, np.pi, (40, 7)).astype(np.float32)
        gi, gd = nn.knn_batch(q, 10)
        oi, od = oracle.knn(0, data, q, 10)
        assert np.array_equal(gi, oi) and np.array_equal(gd, od), step


@pytest.mark.parametrize("dim", [2, 3, 6, 11, 14])
def test_other_dimensions(gpu, oracle, dim):
    rng = np.random.default_rng(dim)
    data = rng.uniform(-5, 5, (3000, dim)).astype(np.float32)
    q = rng.uniform(-5, 5, (64, dim)).astype(np.float32)
    nn = make_nn(gpu, 0, dim, data)
    gi, gd = nn.knn_batch(q, 12)
    oi, od = oracle.knn(0, data, q, 12)
    assert np.array_equal(gi, oi) and np.array_equal(gd, od)


@pytest.mark.parametrize("r,most", [(0.9, 64), (1.2, 512), (3.0, 3000)])
def test_radius_slot_and_segment_paths(gpu, oracle, r, most):
    """Radius results of every path equal the oracle's lists: one pass into
    64-key slots sorted by one wave per query (every query at most 64 hits),
    the second CSR pass sorted by one wave (at most 512), and the second pass
    sorted by the segmented sort (more)."""
    rng = np.random.default_rng(31)
    data = rng.uniform(-1, 1, (3000, 7)).astype(np.float32)
    queries = rng.uniform(-1, 1, (100, 7)).astype(np.float32)
    nn = make_nn(gpu, 0, 7, data, batches=3)
    off, gi, gd = nn.radius_batch(queries, r)
    hits = np.diff(off)
    assert most // 8 < hits.max() <= most
    ores = oracle.radius(0, data, queries, r)
    for j, (oi, od) in enumerate(ores):
        assert np.array_equal(gi[off[j]:off[j + 1]], oi), j
        assert np.array_equal(gd[off[j]:off[j + 1]], od), j


def test_radius_capacity_too_small(gpu, oracle):
    """Outputs smaller than the result count: the call fails with
    OMPL_AMD_ERR_ARG, writes the offsets and the total, and leaves the
    outputs untouched; the Python wrapper then retries at the exact size."""
    import ctypes as C
    import torch
    from ompl_amd import _ptr
    from ompl_amd._lib import load
    rng = np.random.default_rng(32)
    data = rng.uniform(-1, 1, (3000, 7)).astype(np.float32)
    queries = rng.uniform(-1, 1, (50, 7)).astype(np.float32)
    nn = make_nn(gpu, 0, 7, data)
    ores = oracle.radius(0, data, queries, 1.2)
    want = sum(len(oi) for oi, _ in ores)
    qd = torch.from_numpy(queries).cuda()
    off = torch.full((51,), -5, dtype=torch.int64, device="cuda")
    idx = torch.full((8,), -7, dtype=torch.int32, device="cuda")
    dist = torch.full((8,), -7.0, dtype=torch.float32, device="cuda")
    total = C.c_int64()
    code = load().ompl_amd_nn_radius_dev(nn.handle, _ptr(qd), 50, 1.2, _ptr(off), _ptr(idx), _ptr(dist), 8,
                                         C.byref(total), None)
    torch.cuda.synchronize()
    assert code == -1 and total.value == want
    assert off[-1].item() == want and np.array_equal(np.diff(off.cpu().numpy()), [len(oi) for oi, _ in ores])
    assert (idx.cpu() == -7).all() and (dist.cpu() == -7.0).all()
    nn._radius_per_query = 1  # the wrapper's first guess is too small
    off2, gi, gd = nn.radius_dev(qd, 1.2)
    gi, gd, off2 = gi.cpu().numpy(), gd.cpu().numpy(), off2.cpu().numpy()
    for j, (oi, od) in enumerate(ores):
        assert np.array_equal(gi[off2[j]:off2[j + 1]], oi), j
        assert np.array_equal(gd[off2[j]:off2[j + 1]], od), j


@pytest.mark.parametrize("n", [300, 20000])
def test_knn_packed_mixed_requests(gpu, oracle, n):
    """Queries of four requests packed in one launch, each request with its
    own k (1, 10, 33, 64) and some queries with k = 0 or above the batch's
    largest: every query's list equals the oracle's at its k (clamped), in
    the CSR layout; one split (300 points) and many (20,000)."""
    import torch

    rng = np.random.default_rng(40 + n)
    data = rng.uniform(-1, 1, (n, 7)).astype(np.float32)
    nn = make_nn(gpu, 0, 7, data)
    ks = [1, 10, 33, 64]
    sizes = [37, 100, 5, 50]
    q = rng.uniform(-1, 1, (sum(sizes), 7)).astype(np.float32)
    kq = np.concatenate([np.full(m, k, np.int32) for m, k in zip(sizes, ks)])
    kq[3] = 0
    kq[60] = 200  # clamped to the batch's largest
    kmax = 64
    off, gi, gd = nn.knn_packed_dev(torch.from_numpy(q).cuda(), torch.from_numpy(kq).cuda(), kmax)
    off, gi, gd = off.cpu().numpy(), gi.cpu().numpy(), gd.cpu().numpy()
    want_k = np.clip(kq, 0, kmax)
    assert np.array_equal(np.diff(off), want_k)
    for k in sorted(set(want_k.tolist()) - {0}):
        sel = np.nonzero(want_k == k)[0]
        oi, od = oracle.knn(0, data, q[sel], k)
        for j, qi in enumerate(sel):
            assert np.array_equal(gi[off[qi]:off[qi + 1]], oi[j]), (k, qi)
            assert np.array_equal(gd[off[qi]:off[qi + 1]], od[j]), (k, qi)


def test_radius_packed_mixed_requests(gpu, oracle):
    """Queries of three requests packed in one launch, each with its own
    radius, plus a negative and a NaN radius (no results): every query's list
    equals the oracle's at its radius."""
    import torch

    rng = np.random.default_rng(44)
    data = rng.uniform(-1, 1, (5000, 7)).astype(np.float32)
    nn = make_nn(gpu, 0, 7, data)
    radii = [0.4, 0.9, 1.2]
    sizes = [40, 30, 20]
    q = rng.uniform(-1, 1, (sum(sizes), 7)).astype(np.float32)
    rq = np.concatenate([np.full(m, r, np.float32) for m, r in zip(sizes, radii)])
    rq[5] = -1.0
    rq[50] = np.nan
    off, gi, gd = nn.radius_packed_dev(torch.from_numpy(q).cuda(), torch.from_numpy(rq).cuda())
    off, gi, gd = off.cpu().numpy(), gi.cpu().numpy(), gd.cpu().numpy()
    for qi in range(len(q)):
        lo, hi = off[qi], off[qi + 1]
        if not rq[qi] >= 0:
            assert hi == lo, qi
            continue
        (oi, od), = oracle.radius(0, data, q[qi:qi + 1], float(rq[qi]))
        assert np.array_equal(gi[lo:hi], oi), qi
        assert np.array_equal(gd[lo:hi], od), qi
