"""GPU test of ompl_amd_check_motions_noalloc_dev, the batched motion check
whose survivor workspace is a fixed device array of the library: through the
C-ABI, two calls in a row (the second smaller), each equal to
ompl_amd_check_motions_dev (booleans and checked counts bit-exact), and sizes
above the array's limit rejected with an error."""
import ctypes as C
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
MAX_EDGES = 65536


def lvs_config():
    ext = 2 * math.pi * math.sqrt(7)
    return ext * (0.01 / ext)


def noalloc(gpu, env, A, B, lvs, exact):
    import torch

    from ompl_amd._lib import call

    n = A.shape[1]
    out = torch.full((n,), 7, dtype=torch.uint8, device="cuda")
    chk = torch.full((n,), -5, dtype=torch.int32, device="cuda") if exact else None
    call("ompl_amd_check_motions_noalloc_dev", env.handle, C.c_void_p(A.data_ptr()), C.c_void_p(B.data_ptr()), n,
         float(lvs), 1, C.c_void_p(out.data_ptr()), C.c_void_p(chk.data_ptr()) if exact else None,
         C.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return out.cpu().numpy(), (chk.cpu().numpy() if exact else None)


@pytest.mark.parametrize("strategy", [3, 4, 1])
def test_noalloc_matches_allocating_entry_point(gpu, strategy):
    import torch

    s, b = gpu.workload.obstacles(1001, 128, 128)
    env = gpu.Env(s, b)
    lvs = lvs_config()
    if strategy == 4:
        gpu.set_tuning(motion_strategy=4, motion_lanes_per_state=8, motion_obstacle_parts=4)
    else:
        gpu.set_tuning(motion_strategy=strategy, validity_lanes_per_state=7, validity_obstacle_parts=4,
                       motion_lanes_per_state=7, motion_obstacle_parts=9)
    try:
        for n, seed in ((16384, 4), (3001, 9)):  # second call smaller
            a, bb = gpu.workload.edges(seed, n)
            A = torch.from_numpy(a.T.copy()).cuda()
            B = torch.from_numpy(bb.T.copy()).cuda()
            print(f"noalloc strategy {strategy} n {n}", flush=True)
            ref_v, ref_c = gpu.check_motions(env, A, B, lvs, exact_count=True)
            ref_v, ref_c = ref_v.cpu().numpy(), ref_c.cpu().numpy()
            v, c = noalloc(gpu, env, A, B, lvs, exact=True)
            assert np.array_equal(v, ref_v) and np.array_equal(c, ref_c), (n, seed)
            v, _ = noalloc(gpu, env, A, B, lvs, exact=False)
            assert np.array_equal(v, ref_v), (n, seed)
            assert 0 < int(ref_v.sum()) < n
    finally:
        gpu.set_tuning(motion_strategy=3, validity_lanes_per_state=8, validity_obstacle_parts=2,
                       motion_lanes_per_state=8, motion_obstacle_parts=1)


def test_noalloc_with_environment_shapes_many_survivors(gpu):
    """A sparse scene with its own launch shapes (about 10,000 survivors, the
    walk's chunk-major-only form): the fixed-array entry point uses the
    environment's shapes and matches the allocating one, twice in a row."""
    import torch

    s, b = gpu.workload.obstacles(1002, 16, 16)
    env = gpu.Env(s, b)
    env.set_tuning(motion_strategy=3, validity_lanes_per_state=4, validity_obstacle_parts=1,
                   motion_lanes_per_state=7, motion_obstacle_parts=1, motion_waves_per_cu=32)
    lvs = lvs_config()
    for n, seed in ((16384, 9), (5000, 10)):
        a, bb = gpu.workload.edges(seed, n, -0.5, 1.0, 0.05, 5.0)
        A = torch.from_numpy(a.T.copy()).cuda()
        B = torch.from_numpy(bb.T.copy()).cuda()
        print(f"noalloc environment shapes n {n}", flush=True)
        ref_v, ref_c = gpu.check_motions(env, A, B, lvs, exact_count=True)
        ref_v, ref_c = ref_v.cpu().numpy(), ref_c.cpu().numpy()
        v, c = noalloc(gpu, env, A, B, lvs, exact=True)
        assert np.array_equal(v, ref_v) and np.array_equal(c, ref_c), n
        v, _ = noalloc(gpu, env, A, B, lvs, exact=False)
        assert np.array_equal(v, ref_v), n
    assert int(ref_v.sum()) > 0
    env.close()


def test_noalloc_rejects_batches_above_its_array(gpu):
    import torch

    from ompl_amd._lib import OmplAmdError

    s, b = gpu.workload.obstacles(1001, 16, 16)
    env = gpu.Env(s, b)
    n = MAX_EDGES + 1
    A = torch.zeros((7, n), dtype=torch.float64, device="cuda")
    with pytest.raises(OmplAmdError, match="at most 65536"):
        noalloc(gpu, env, A, A, lvs_config(), exact=False)
