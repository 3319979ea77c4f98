"""GPU tests of the simulator boundary: ops that only decompose, sub-circuits, initial-state
validation, MemoryError, amplitudes and JSON-loaded circuits, each against the oracle or the
reference tests' known answers."""

import os
import sys

import numpy as np
import pytest
import sympy

import cirq_amd as cirq
from oracle import cirq_np as O
from tests.helpers import fidelity, to_oracle_ops

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class _Bell(cirq.Gate):
    """A gate with no matrix of its own: only a decomposition."""

    def _num_qubits_(self):
        return 2

    def _decompose_(self, qubits):
        a, b = qubits
        yield cirq.H(a)
        yield cirq.CNOT(a, b)
        yield cirq.rz(0.3).on(b)


class _Ladder(cirq.Gate):
    """Six qubits: wider than a fused block, applied through its decomposition."""

    def _num_qubits_(self):
        return 6

    def _decompose_(self, qubits):
        for i, q in enumerate(qubits):
            yield cirq.ry(0.2 + 0.1 * i).on(q)
        for a, b in zip(qubits[:-1], qubits[1:]):
            yield cirq.CZ(a, b) ** 0.7
        yield _Bell().on(qubits[0], qubits[5])


def _expanded_oracle_state(c, qs, n):
    from cirq_amd.circuits import expand_for_simulation

    return O.simulate_state_vector(to_oracle_ops(expand_for_simulation(c), qs), n)


def test_decompose_only_gates_match_oracle(gpu):
    _log("decompose-only gates")
    qs = cirq.LineQubit.range(8)
    c = cirq.Circuit(cirq.H.on_each(*qs), _Bell().on(qs[1], qs[6]), _Ladder().on(*qs[1:7]),
                     cirq.X(qs[0]) ** 0.3, _Ladder().on(qs[7], qs[0], qs[2], qs[4], qs[3], qs[5]))
    got = cirq.Simulator().simulate(c, qubit_order=qs).final_state_vector
    want = _expanded_oracle_state(c, qs, 8)
    assert np.max(np.abs(got - want)) <= 1e-5 and fidelity(got, want) >= 1 - 1e-6
    # the hand-expanded circuit is the same state
    hand = cirq.Circuit(cirq.H.on_each(*qs), cirq.H(qs[1]), cirq.CNOT(qs[1], qs[6]), cirq.rz(0.3).on(qs[6]))
    psi = O.simulate_state_vector(to_oracle_ops(hand, qs), 8)
    psi_g = cirq.Simulator().simulate(cirq.Circuit(cirq.H.on_each(*qs), _Bell().on(qs[1], qs[6])),
                                      qubit_order=qs).final_state_vector
    assert np.max(np.abs(psi_g - psi)) <= 1e-5


def test_circuit_operation_matches_oracle(gpu):
    _log("CircuitOperation")
    qs = cirq.LineQubit.range(6)
    t = sympy.Symbol("t")
    sub = cirq.FrozenCircuit(cirq.ry(t).on(qs[0]), cirq.CZ(qs[0], qs[1]), cirq.X(qs[1]) ** 0.25)
    op = cirq.CircuitOperation(sub, repetitions=3, qubit_map={qs[0]: qs[4], qs[1]: qs[2]}, param_resolver={"t": 0.7})
    nested = cirq.CircuitOperation(cirq.FrozenCircuit(op, cirq.H(qs[5])), repetitions=2)
    c = cirq.Circuit(cirq.H.on_each(*qs), op, cirq.CNOT(qs[2], qs[3]), nested,
                     cirq.CircuitOperation(sub, repetitions=-1, param_resolver={"t": sympy.Symbol("u")}))
    r = cirq.ParamResolver({"u": -0.4})
    got = cirq.Simulator().simulate(c, r, qubit_order=qs).final_state_vector
    from cirq_amd.circuits import expand_for_simulation

    want = O.simulate_state_vector(to_oracle_ops(expand_for_simulation(c._resolve_parameters_(r)), qs), 6)
    assert np.max(np.abs(got - want)) <= 1e-5 and fidelity(got, want) >= 1 - 1e-6


def test_circuit_operation_measurement_keys(gpu):
    qs = cirq.LineQubit.range(2)
    sub = cirq.FrozenCircuit(cirq.X(qs[0]), cirq.measure(qs[0], key="m"))
    c = cirq.Circuit(cirq.CircuitOperation(sub, repetitions=2, use_repetition_ids=True, qubit_map={qs[0]: qs[1]}))
    res = cirq.Simulator(seed=1).run(c, repetitions=5)
    assert set(res.measurements) == {"0:m", "1:m"}
    assert res.measurements["0:m"].tolist() == [[1]] * 5 and res.measurements["1:m"].tolist() == [[0]] * 5


def test_initial_states_validated_as_cirq_does(gpu):
    _log("initial-state validation")
    qs = cirq.LineQubit.range(3)
    c = cirq.Circuit(cirq.H(qs[0]), cirq.CNOT(qs[0], qs[2]), cirq.ry(0.4).on(qs[1]))
    sim = cirq.Simulator()
    with pytest.raises(ValueError, match="not normalized"):
        sim.simulate(c, initial_state=np.full(8, 3.0, dtype=np.complex64))  # |a| > 1: would overflow FP16 blocks
    with pytest.raises(ValueError, match="incorrect size"):
        sim.simulate(c, initial_state=np.ones(4, dtype=np.complex64) / 2)
    rng = np.random.RandomState(3)
    v = rng.randn(8) + 1j * rng.randn(8)
    v = (v / np.linalg.norm(v)).astype(np.complex64)
    got = sim.simulate(c, initial_state=v).final_state_vector
    want = O.simulate_state_vector(to_oracle_ops(c, qs), 3, initial_state=0)
    U = np.stack([O.simulate_state_vector(to_oracle_ops(c, qs), 3, initial_state=b) for b in range(8)], axis=1)
    want = U @ v.astype(np.complex128)
    assert np.max(np.abs(got - want)) <= 1e-5
    # per-qubit basis values name a basis state
    got = sim.simulate(c, initial_state=np.array([1, 0, 1])).final_state_vector
    assert np.max(np.abs(got - U[:, 5])) <= 1e-5
    dsim = cirq.DensityMatrixSimulator()
    with pytest.raises(ValueError, match="trace 1"):
        dsim.simulate(c, initial_state=2 * np.eye(8, dtype=np.complex64) / 8)
    with pytest.raises(ValueError, match="Hermitian"):
        dsim.simulate(c, initial_state=np.eye(8, dtype=np.complex64) / 8 + 0.1j * np.triu(np.ones((8, 8)), 1))
    rho = dsim.simulate(c, initial_state=np.eye(8, dtype=np.complex64) / 8).final_density_matrix
    assert np.max(np.abs(rho - U @ (np.eye(8) / 8) @ U.conj().T)) <= 1e-6


def test_memory_error_for_states_the_device_cannot_hold(gpu):
    # cirq-core/cirq/sim/density_matrix_simulator_test.py:1502-1512: 59 qubits, entangled state
    circuit = cirq.Circuit()
    for i in range(59):
        for _ in range(9):
            circuit.append(cirq.X(cirq.LineQubit(i)))
        circuit.append(cirq.measure(cirq.LineQubit(i)))
    with pytest.raises(MemoryError, match="Unable to allocate"):
        cirq.DensityMatrixSimulator(split_untangled_states=False).simulate(circuit)
    with pytest.raises(MemoryError, match="Unable to allocate"):
        cirq.Simulator().simulate(cirq.Circuit(cirq.H.on_each(*cirq.LineQubit.range(40))))


def test_compute_amplitudes_known_answers_and_oracle(gpu):
    # cirq-core/cirq/sim/sparse_simulator_test.py:883-905
    a, b = cirq.LineQubit.range(2)
    c = cirq.Circuit(cirq.X(a), cirq.H(a), cirq.H(b))
    sim = cirq.Simulator()
    np.testing.assert_allclose(np.array(sim.compute_amplitudes(c, [0])), np.array([0.5]), atol=1e-7)
    np.testing.assert_allclose(np.array(sim.compute_amplitudes(c, [1, 2, 3])), [0.5, -0.5, -0.5], atol=1e-7)
    np.testing.assert_allclose(np.array(sim.compute_amplitudes(c, (1, 2, 3), qubit_order=(b, a))), [-0.5, 0.5, -0.5],
                               atol=1e-7)
    with pytest.raises(ValueError, match="1-dimensional"):
        sim.compute_amplitudes(c, np.array([[0, 0]]))
    # a parameter sweep on 12 qubits against the oracle
    from cirq_amd import generators

    circ, syms, vals = generators.hardware_efficient_ansatz(12, 3, seed=2)
    qs = sorted(circ.all_qubits())
    rng = np.random.RandomState(5)
    bits = rng.randint(0, 2**12, 50).tolist()
    resolvers = [cirq.ParamResolver(dict(zip(syms, rng.uniform(-3, 3, len(syms))))) for _ in range(3)]
    got = sim.compute_amplitudes_sweep(circ, bits, resolvers)
    for r, amps in zip(resolvers, got):
        want = O.simulate_state_vector(to_oracle_ops(circ, qs, resolver=r), 12)[bits]
        assert np.max(np.abs(np.array(amps) - want)) <= 1e-5


def test_json_loaded_circuit_on_the_gpu(gpu):
    """The reference's own Circuit.json fixture, loaded and simulated on the HIP path."""
    circuits = cirq.read_json(os.path.join(GOLDEN, "json", "Circuit.json"))
    assert len(circuits) == 3
    # H on 5 qubits then measure all, one joint factor: NumPy's choice on the oracle's probabilities
    c0 = circuits[0]
    qs = sorted(c0.all_qubits())
    got = cirq.Simulator(seed=1234, split_untangled_states=False).run(c0, repetitions=200).measurements["0,1,2,3,4"]
    psi = O.simulate_state_vector(to_oracle_ops(c0, qs), 5)
    want = O.sample_state_vector(psi, list(range(5)), 200, np.random.RandomState(1234))
    assert np.array_equal(got, want)
    # TOFFOLI then X**0.123; X**theta with theta resolved
    for c, r in ((circuits[1], None), (circuits[2], cirq.ParamResolver({"theta": 0.37}))):
        qs = sorted(c.all_qubits())
        got = cirq.Simulator().simulate(c, r, qubit_order=qs).final_state_vector
        want = O.simulate_state_vector(to_oracle_ops(c, qs, resolver=r), len(qs))
        _log(f"Circuit.json: {len(qs)} qubits, max|err| {np.max(np.abs(got - want)):.2e}")
        assert np.max(np.abs(got - want)) <= 1e-5 and fidelity(got, want) >= 1 - 1e-6


def test_expectation_values_known_answers_and_measurements(gpu):
    # cirq-core/cirq/sim/sparse_simulator_test.py:630-708
    q0, q1, q2 = cirq.LineQubit.range(3)
    sim = cirq.Simulator()
    psum1 = cirq.Z(q0) + 3.2 * cirq.Z(q1)
    psum2 = -1 * cirq.X(q0) + 2 * cirq.X(q1)
    r = sim.simulate_expectation_values(cirq.Circuit(cirq.I(q0), cirq.X(q1)), [psum1, psum2])
    assert abs(r[0] - -2.2) < 1e-6 and abs(r[1]) < 1e-6
    r = sim.simulate_expectation_values(cirq.Circuit(cirq.H(q0), cirq.H(q1)), [psum1, psum2])
    assert abs(r[0]) < 1e-6 and abs(r[1] - 1) < 1e-6
    r = sim.simulate_expectation_values(cirq.Circuit(cirq.I(q0), cirq.H(q1)), cirq.Z(q0) + cirq.X(q1))
    assert abs(r[0] - 2) < 1e-6
    c = cirq.Circuit(cirq.H(q0), cirq.H(q1), cirq.X(q2))
    obs = cirq.X(q0) + cirq.X(q1) - cirq.Z(q2)
    assert abs(sim.simulate_expectation_values(c, obs)[0] - 3) < 1e-6
    assert abs(sim.simulate_expectation_values(c, obs, qubit_order=[q1, q2, q0])[0] - 3) < 1e-6
    # terminal measurement: refused unless permitted, then the collapsed state's value (+-1)
    term = cirq.Circuit(cirq.H(q0), cirq.measure(q0))
    with pytest.raises(ValueError):
        sim.simulate_expectation_values(term, cirq.Z(q0))
    seen = {round(float(np.real(sim.simulate_expectation_values(term, cirq.Z(q0), permit_terminal_measurements=True)[0])))
            for _ in range(40)}
    assert seen == {-1, 1}
    # a mid-circuit measurement is not terminal: one sampled trajectory, the same as simulate's
    mid = cirq.Circuit(cirq.H(q0), cirq.measure(q0, key="m"), cirq.H(q0), cirq.CNOT(q0, q1), cirq.ry(0.4).on(q1))
    for seed in range(4):
        ev = cirq.Simulator(seed=seed).simulate_expectation_values(mid, [cirq.Z(q0) * cirq.Z(q1), cirq.X(q1)])
        psi = cirq.Simulator(seed=seed).simulate(mid).final_state_vector.astype(np.complex128)
        qs = [q0, q1]
        want = [sum(O.expectation_from_state_vector(psi, t, 2) for t in
                    [(complex(x.coefficient), {qs.index(q): p._name for q, p in x.items()}) for x in cirq.PauliSum.wrap(o)])
                for o in (cirq.Z(q0) * cirq.Z(q1), cirq.X(q1))]
        np.testing.assert_allclose(np.array(ev, dtype=complex), want, atol=1e-6)
