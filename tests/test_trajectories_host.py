"""CPU tests of the batched-trajectory host logic: NumPy's draws restated row by row.

Each stochastic op of a repetition consumes one uniform of the RandomState, so the batched runner
draws all of them up front and turns each row into the outcome NumPy's RandomState.choice (or the
Kraus walk of cirq-core/cirq/sim/state_vector_simulation_state.py:224-250) would give for it."""

import numpy as np
import pytest

from cirq_amd.trajectories import TrajectoryBatch, choice_rows, kraus_draws


@pytest.mark.parametrize("seed", range(5))
def test_choice_rows_matches_randomstate_choice(seed):
    rng = np.random.RandomState(100 + seed)
    R, K = 300, int(rng.randint(1, 17))
    p = rng.exponential(size=(R, K))
    p[rng.uniform(size=(R, K)) < 0.2] = 0.0
    p[:, 0] += 1e-3  # no all-zero row
    p /= p.sum(axis=1, keepdims=True)
    a = np.random.RandomState(seed)
    want = np.array([a.choice(K, p=row) for row in p])
    b = np.random.RandomState(seed)
    got = choice_rows(p, b.random_sample(R))
    assert np.array_equal(got, want)
    assert a.random_sample() == b.random_sample()  # the same number of draws was consumed


def _kraus_walk(weights_row, p):
    """The walk one repetition makes (the simulator's one-state channel path, which restates
    cirq-core/cirq/sim/state_vector_simulation_state.py:224-250)."""
    fallback_weight, fallback_index = 0.0, 0
    weight = None
    index = -1
    for index, w in enumerate(weights_row):
        weight = w
        if weight > fallback_weight:
            fallback_index, fallback_weight = index, weight
        p -= weight
        if p < 0:
            break
    if p >= 0 or weight == 0:
        return fallback_index, fallback_weight
    return index, weight


@pytest.mark.parametrize("seed", range(5))
def test_kraus_draws_match_the_walk(seed):
    rng = np.random.RandomState(seed)
    R, M = 500, int(rng.randint(1, 9))
    W = rng.exponential(size=(R, M))
    W[rng.uniform(size=W.shape) < 0.25] = 0.0
    W /= W.sum(axis=1, keepdims=True) + rng.uniform(-0.05, 0.05, size=(R, 1))  # sums near, not at, 1
    W[:5] = 0.0  # rows whose walk never ends: fallback to operator 0 with weight 0
    u = rng.uniform(size=R)
    sel, w = kraus_draws(W, u)
    for r in range(R):
        i, wi = _kraus_walk(W[r], u[r])
        assert sel[r] == i and w[r] == wi, r


def test_kraus_draws_amplitude_damping_closed_form():
    """From |1>, amplitude damping's weights are (1 - gamma, gamma): operator 0 exactly when
    u < 1 - gamma."""
    g = 0.3
    u = np.linspace(0, 0.999, 1000)
    W = np.tile([1 - g, g], (len(u), 1))
    sel, w = kraus_draws(W, u)
    assert np.array_equal(sel, (u >= 1 - g).astype(int))
    assert np.allclose(w, np.where(u < 1 - g, 1 - g, g))


def test_supported_ops():
    import cirq_amd as cirq

    q = cirq.LineQubit.range(4)
    acts = [("op", cirq.measure(*q)), ("op", cirq.depolarize(0.1)(q[0])), ("op", cirq.amplitude_damp(0.2)(q[1]))]
    assert TrajectoryBatch.supported(acts, dm=False)
    assert TrajectoryBatch.supported([("op", cirq.measure(q[0]))], dm=True)
