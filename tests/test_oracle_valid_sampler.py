"""CPU pins of the oracle's UniformValidStateSampler::sample restatement: the
calls consume the generator's uniform draws in order (replayed with the
mt19937 sampler, itself pinned to std::mt19937 golden output), each call ends
at its first valid draw or after `attempts` draws, and found/draws agree with
brute-force validity of the replayed draws."""
import numpy as np
import pytest


@pytest.mark.parametrize("n_obs,attempts", [(16, 100), (64, 3), (256, 5)])
def test_valid_sampler_replays_uniform_draws(oracle, n_obs, attempts):
    from ompl_amd import workload  # host-side generator, no device needed

    s, b = workload.obstacles(1002, n_obs, n_obs)
    env = oracle.OracleEnv(s, b)
    n_calls = 40
    states, found, draws = oracle.Mt19937(21).sample_valid(env, n_calls, attempts)
    total = int(draws.sum())
    lo, hi = np.full(7, -np.pi), np.full(7, np.pi)
    stream = oracle.Mt19937(21).sample_uniform(lo, hi, total)
    valid = np.array([oracle.state_valid(env, r.astype(np.float32)) for r in stream])
    pos = 0
    for c in range(n_calls):
        d = int(draws[c])
        assert 1 <= d <= attempts
        seg = valid[pos:pos + d]
        assert not seg[:-1].any()  # every earlier draw of the call was invalid
        assert bool(found[c]) == bool(seg[-1])
        if not found[c]:
            assert d == attempts
        assert np.array_equal(states[c], stream[pos + d - 1])
        pos += d
    assert 0 < found.sum() or n_obs == 256
