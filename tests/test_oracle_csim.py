"""The C complex128 kernels (oracle/csim.c) against the NumPy restatement of Cirq's algorithm.

The C kernels write the fixtures the large-size GPU tests compare with (28-qubit grid circuit,
12-qubit noisy density matrix), so they are pinned here on random circuits small enough for the
NumPy oracle: every gate family and channel the tests use, targets in every order.
"""

import numpy as np
import pytest

from cirq_amd import generators
from oracle import cirq_np as O
from oracle import csim
from tests.helpers import to_oracle_ops
from tests.test_host import _random_circuit


@pytest.mark.parametrize("seed", range(3))
def test_state_vector_matches_numpy_oracle(seed):
    rng = np.random.RandomState(40 + seed)
    c, qs = _random_circuit(9, 150, rng)
    ops = to_oracle_ops(c, qs)
    want = O.simulate_state_vector(ops, 9)
    got = csim.simulate_state_vector(ops, 9)
    assert np.max(np.abs(got - want)) < 1e-12


def test_grid_circuit_matches_numpy_oracle():
    c = generators.grid_circuit(3, 4, 12, seed=2)
    qs = sorted(c.all_qubits())
    ops = to_oracle_ops(c, qs)
    assert np.max(np.abs(csim.simulate_state_vector(ops, 12) - O.simulate_state_vector(ops, 12))) < 1e-12


@pytest.mark.parametrize("seed", range(3))
def test_density_matrix_matches_numpy_oracle(seed):
    rng = np.random.RandomState(70 + seed)
    c, qs = _random_circuit(5, 60, rng, with_channels=True)
    ops = to_oracle_ops(c, qs)
    want = O.simulate_density_matrix(ops, 5).reshape(-1)
    got = csim.simulate_density_matrix(ops, 5)
    assert np.max(np.abs(got - want)) < 1e-12


def test_noisy_line_matches_numpy_oracle():
    nc = generators.noisy_line_circuit(5, 6, seed=1, p=0.05, gamma=0.04)
    qs = sorted(nc.all_qubits())
    ops = to_oracle_ops(nc, qs)
    want = O.simulate_density_matrix(ops, 5).reshape(-1)
    assert np.max(np.abs(csim.simulate_density_matrix(ops, 5) - want)) < 1e-12


def test_expect_z_matches_numpy():
    rng = np.random.RandomState(1)
    psi = rng.randn(2**10) + 1j * rng.randn(2**10)
    masks = [0, 1, 0b1000000001, 0b0110, 2**10 - 1]
    got = csim.expect_z(psi, 10, masks)
    p = np.abs(psi) ** 2
    idx = np.arange(2**10)
    for g, z in zip(got, masks):
        sign = 1 - 2 * (np.array([bin(v).count("1") for v in idx & z]) & 1)
        assert abs(g - np.sum(sign * p)) < 1e-9


def test_rejects_repeated_or_out_of_range_axes():
    psi = np.zeros(8, dtype=np.complex128)
    with pytest.raises(ValueError):
        csim.apply_matrix(psi, 3, np.eye(4), [1, 1])
    with pytest.raises(ValueError):
        csim.apply_matrix(psi, 3, np.eye(4), [0, 3])
