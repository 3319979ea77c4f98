"""GPU tests of Simulator.sample_expectation_values: eigenstates exactly, a random circuit within
6 standard errors of the exact expectation values (simulate_expectation_values, itself checked
against the oracle elsewhere), reproducible with a seed, and on the density-matrix simulator."""

import sys

import numpy as np
import pytest
import sympy

import cirq_amd as cirq
from cirq_amd import generators

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_eigenstates_exact(gpu):
    _log("sample_expectation_values eigenstates")
    q = cirq.LineQubit.range(3)
    c = cirq.Circuit(cirq.X(q[0]), cirq.H(q[1]), cirq.rx(np.pi / 2).on(q[2]))
    res = cirq.Simulator(seed=1).sample_expectation_values(
        c, [cirq.Z(q[0]), cirq.X(q[1]) * cirq.Z(q[0]), cirq.Y(q[2]) + 0.5 * cirq.X(q[1])], num_samples=3000)
    np.testing.assert_allclose(res, [[-1.0, -1.0, -0.5]], atol=1e-12)


def test_random_circuit_statistics_and_seed(gpu):
    _log("sample_expectation_values 10q random circuit")
    c = generators.grid_circuit(2, 5, 8, seed=2)
    qs = sorted(c.all_qubits())
    obs = [generators.random_pauli_sum(qs, 6, 3, seed=s, include_z=False) for s in range(3)]
    N = 30_000
    res = cirq.Simulator(seed=7).sample_expectation_values(c, obs, num_samples=N)
    exact = cirq.Simulator().simulate_expectation_values(c, obs)
    for j, s in enumerate(obs):
        bound = 6 * sum(abs(t.coefficient) for t in s) / np.sqrt(N)
        assert abs(res[0][j] - exact[j].real) < bound, (j, res[0][j], exact[j])
    again = cirq.Simulator(seed=7).sample_expectation_values(c, obs, num_samples=N)
    assert again == res


def test_sweep_and_density_matrix(gpu):
    _log("sample_expectation_values sweep + density matrix")
    q = cirq.LineQubit.range(2)
    t = sympy.Symbol("t")
    c = cirq.Circuit(cirq.rx(t).on(q[0]), cirq.CNOT(q[0], q[1]))
    res = cirq.Simulator(seed=3).sample_expectation_values(
        c, [cirq.Z(q[0]) * cirq.Z(q[1]), cirq.Z(q[1])], num_samples=4000, params=cirq.Points("t", [0.0, np.pi]))
    np.testing.assert_allclose(res, [[1, 1], [1, -1]], atol=1e-12)
    noisy = cirq.Circuit(cirq.H(q[0]), cirq.amplitude_damp(0.3).on(q[0]), cirq.depolarize(0.1).on(q[1]))
    dm = cirq.DensityMatrixSimulator(seed=5)
    N = 40_000
    got = dm.sample_expectation_values(noisy, [cirq.Z(q[0]), cirq.X(q[0]), cirq.Z(q[1])], num_samples=N)
    exact = dm.simulate_expectation_values(noisy, [cirq.Z(q[0]), cirq.X(q[0]), cirq.Z(q[1])])
    for j in range(3):
        assert abs(got[0][j] - exact[j].real) < 6 / np.sqrt(N), (j, got[0][j], exact[j])
