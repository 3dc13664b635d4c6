"""GPU parity of SpaceInformation::checkMotion(states, count[,
firstInvalidStateIndex]) over precomputed state lists (batched and plain
single call) against the CPU oracle: identical validity and first invalid
index, including empty, single-state and clamped lists."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lists_scene(gpu, oracle):
    s, b = gpu.workload.obstacles(1001, 32, 32)
    env = gpu.Env(s, b)
    oenv = oracle.OracleEnv(s, b)
    q = gpu.workload.configs(5, 4000)
    v = np.array([oracle.state_valid(oenv, r.astype(np.float32)) for r in q])
    return env, oenv, q[v], q[~v]


def make_lists(good, bad, seed, n_lists, max_count):
    rng = np.random.default_rng(seed)
    rows, offsets = [], [0]
    for i in range(n_lists):
        count = int(rng.integers(0, max_count + 1)) if i % 7 else i % 3  # some 0/1/2-state lists
        st = good[rng.integers(0, len(good), count)].copy()
        if count and rng.random() < 0.5:
            k = int(rng.integers(1, 4))
            st[rng.integers(0, count, k)] = bad[rng.integers(0, len(bad), k)]
        rows.append(st)
        offsets.append(offsets[-1] + count)
    return np.concatenate(rows) if rows else np.zeros((0, 7)), np.array(offsets, np.int64)


# launch shapes of the list walk (the motion walk's tuning fields)
SHAPES = [(8, 1), (7, 1), (7, 4), (7, 9), (1, 1), (8, 8)]


@pytest.fixture
def walk_shape(gpu, request):
    saved = gpu.get_tuning()
    lps, ops = request.param
    gpu.set_tuning(motion_lanes_per_state=lps, motion_obstacle_parts=ops)
    yield request.param
    gpu.set_tuning(**saved)


@pytest.mark.parametrize("walk_shape", SHAPES, indirect=True)
@pytest.mark.parametrize("seed,n_lists,max_count", [(1, 50, 10), (2, 400, 600), (3, 3000, 40)])
def test_state_lists_match_oracle(gpu, oracle, lists_scene, seed, n_lists, max_count, walk_shape):
    import torch

    env, oenv, good, bad = lists_scene
    states, off = make_lists(good, bad, seed, n_lists, max_count)
    print(f"state lists seed {seed}: {n_lists} lists, {states.shape[0]} states, shape {walk_shape}", flush=True)
    q = torch.from_numpy(states.T.copy()).cuda()
    o = torch.from_numpy(off).cuda()
    valid, first = gpu.check_state_lists(env, q, o)
    valid, first = valid.cpu().numpy(), first.cpu().numpy()
    for i in range(n_lists):
        st = states[off[i]:off[i + 1]]
        ok, _ = oracle.check_states(oenv, st)
        ok2, f = oracle.check_states_first_invalid(oenv, st)
        assert bool(valid[i]) == ok == ok2, i
        assert first[i] == (-1 if ok else f), i


def test_state_lists_clamped_offsets(gpu, oracle, lists_scene):
    import torch

    env, oenv, good, bad = lists_scene
    states = good[:20].copy()
    states[12] = bad[0]
    q = torch.from_numpy(states.T.copy()).cuda()
    # [-5, 10): clamped to [0, 10); [10, 99): clamped to [10, 20); [15, 3): empty
    o = torch.tensor([-5, 10, 99, 3], dtype=torch.int64).cuda()
    valid, first = gpu.check_state_lists(env, q, o)
    assert valid.cpu().tolist() == [1, 0, 1]
    assert first.cpu().tolist() == [-1, 2, -1]


@pytest.mark.parametrize("walk_shape", [(8, 1), (7, 9)], indirect=True)
def test_check_states_single_call(gpu, oracle, lists_scene, walk_shape):
    env, oenv, good, bad = lists_scene
    rng = np.random.default_rng(5)
    for count in (0, 1, 2, 3, 9, 64, 65, 300):
        st = good[rng.integers(0, len(good), count)].copy()
        for plant in (None, 0, count // 2, count - 1):
            s2 = st.copy()
            if plant is not None and count:
                s2[plant] = bad[1]
            ok, f = gpu.check_states(env, s2)
            ook, of = oracle.check_states_first_invalid(oenv, s2)
            assert ok == ook and f == of, (count, plant)
