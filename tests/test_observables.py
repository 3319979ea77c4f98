"""CPU tests of the sampled-expectation machinery (cirq_amd.observables): greedy grouping with
Cirq's dict semantics, basis-change parameters, chunking, aggregation per PauliSum and sweep point.
The sampler here is the oracle's exact complex128 simulation sampled with NumPy, so the host logic
is checked against exact expectation values (statistically) and against eigenstates (exactly)."""

import warnings

import numpy as np
import pytest
import sympy

import cirq_amd as cirq
from cirq_amd import observables as obs
from cirq_amd.study import ResultDict, to_resolvers
from oracle import cirq_np as O
from tests.helpers import terms_of, to_oracle_ops


class OracleSampler:
    """run_sweep on the oracle: exact state, NumPy choice over all measured qubits."""

    def __init__(self, seed=0):
        self.rng = np.random.RandomState(seed)
        self.calls = []

    def run_sweep(self, program, params, repetitions):
        qubits = sorted(program.all_qubits())
        (mop,) = [op for op in program.all_operations() if type(getattr(op, "gate", None)).__name__ == "MeasurementGate"]
        out = []
        for r in to_resolvers(params):
            psi = O.simulate_state_vector(to_oracle_ops(program, qubits, r), len(qubits))
            axes = [qubits.index(q) for q in mop.qubits]
            bits = O.sample_state_vector(psi, axes, repetitions, self.rng)
            out.append(ResultDict(params=r, measurements={"z": bits.astype(np.uint8)}))
        self.calls.append((len(out), repetitions))
        return out


def _exact(circuit, psum, resolver=None):
    qs = sorted(circuit.all_qubits() | set(cirq.PauliSum.wrap(psum).qubits))
    psi = O.simulate_state_vector(to_oracle_ops(circuit, qs, resolver), len(qs))
    return sum(O.expectation_from_state_vector(psi, t, len(qs)) for t in terms_of(cirq.PauliSum.wrap(psum), qs)).real


def test_greedy_grouping_follows_dict_semantics():
    q = cirq.LineQubit.range(2)
    X0, Z1, Y0, X1 = cirq.X(q[0]), cirq.Z(q[1]), cirq.Y(q[0]), cirq.X(q[1])
    strings = [X0, Z1, X0 * Z1, Y0, cirq.Z(q[0]) * Z1, X1]
    settings = list(obs.observables_to_settings(strings, q))
    grouped = obs.group_settings_greedy(settings)
    keys = [dict(k.observable.items()) for k in grouped]
    assert keys == [{q[0]: cirq.X, q[1]: cirq.Z}, {q[0]: cirq.Z, q[1]: cirq.Z}, {q[0]: cirq.Y, q[1]: cirq.X}]
    members = [[s.observable for s in v] for v in grouped.values()]
    assert members == [[X0, Z1, X0 * Z1], [cirq.Z(q[0]) * Z1], [Y0, X1]]
    assert all(k.observable.coefficient == 1 for k in grouped)


def test_basis_change_parameters():
    q = cirq.LineQubit.range(3)
    s = obs.InitObsSetting(obs.zeros_state(q), cirq.X(q[0]) * cirq.Y(q[2]))
    p = obs._params_for_setting(s, [False] * 3, q, needs_init_layer=False)
    assert p == {"q(0)-Xf": 0, "q(0)-Yf": -0.5, "q(1)-Xf": 0, "q(1)-Yf": 0, "q(2)-Xf": 0.5, "q(2)-Yf": 0}
    p = obs._params_for_setting(s, [True] * 3, q, needs_init_layer=True)
    assert p["q(0)-Yf"] == 0.5 and p["q(1)-Xf"] == 1 and p["q(2)-Xf"] == -0.5 and p["q(1)-Xi"] == 0


def test_eigenstates_are_exact():
    q = cirq.LineQubit.range(3)
    c = cirq.Circuit(cirq.X(q[0]), cirq.H(q[1]), cirq.rx(np.pi / 2).on(q[2]))
    res = obs.sample_expectation_values(
        OracleSampler(), c, [cirq.Z(q[0]), cirq.X(q[1]), cirq.Z(q[0]) * cirq.X(q[1]), 2.0 * cirq.Y(q[2]) * cirq.X(q[1]),
                             cirq.Y(q[2]) + 0.5 * cirq.X(q[1])], num_samples=500)
    np.testing.assert_allclose(res, [[-1.0, 1.0, -1.0, -2.0, -0.5]], atol=1e-12)


def test_string_shared_by_two_sums_counts_in_the_last():
    """As in the reference, each PauliString maps to the last sum holding it
    (cirq-core/cirq/work/sampler.py:366-372 builds that map with a dict): a string in two sums is
    measured twice and both means land in the later sum."""
    q = cirq.LineQubit.range(2)
    c = cirq.Circuit(cirq.X(q[0]), cirq.H(q[1]))
    res = obs.sample_expectation_values(OracleSampler(), c, [cirq.Z(q[0]), cirq.Z(q[0]) + 0.5 * cirq.X(q[1])],
                                        num_samples=100)
    np.testing.assert_allclose(res, [[0.0, -1.5]], atol=1e-12)


def test_statistics_match_exact_values_and_sums_are_aggregated():
    q = cirq.LineQubit.range(3)
    c = cirq.Circuit(cirq.H(q[0]), cirq.CNOT(q[0], q[1]), cirq.ry(0.7).on(q[2]), cirq.rz(0.3).on(q[1]),
                     cirq.CZ(q[1], q[2]), cirq.rx(0.4).on(q[0]))
    sums = [cirq.Z(q[0]) * cirq.Z(q[1]) + 0.5 * cirq.X(q[2]), cirq.X(q[0]) * cirq.X(q[1]) - 0.3 * cirq.Y(q[1]),
            cirq.PauliSum.wrap(cirq.Z(q[2]))]
    N = 40_000
    res = obs.sample_expectation_values(OracleSampler(1), c, sums, num_samples=N)
    for j, s in enumerate(sums):
        bound = 6 * sum(abs(t.coefficient) for t in s) / np.sqrt(N)
        assert abs(res[0][j] - _exact(c, s)) < bound, (j, res[0][j], _exact(c, s))


def test_parameter_sweep_points():
    q = cirq.LineQubit.range(2)
    t = sympy.Symbol("t")
    c = cirq.Circuit(cirq.rx(t).on(q[0]), cirq.CNOT(q[0], q[1]))
    sweep = cirq.Points("t", [0.0, np.pi, np.pi / 2])
    res = obs.sample_expectation_values(OracleSampler(2), c, [cirq.Z(q[0]), cirq.Z(q[0]) * cirq.Z(q[1])],
                                        num_samples=20_000, params=sweep)
    assert len(res) == 3
    np.testing.assert_allclose(res[0], [1, 1], atol=1e-12)
    np.testing.assert_allclose(res[1], [-1, 1], atol=1e-12)
    assert abs(res[2][0]) < 6 / np.sqrt(20_000) and res[2][1] == 1


def test_chunking_and_readout_symmetrization():
    q = cirq.LineQubit.range(2)
    c = cirq.Circuit(cirq.H(q[0]), cirq.X(q[1]))
    s = OracleSampler(3)
    res = obs.measure_observables(c, [cirq.X(q[0]), cirq.Z(q[1]), cirq.Y(q[0])], s,
                                  obs.RepetitionsStoppingCriteria(25_000, repetitions_per_chunk=10_000))
    assert s.calls == [(2, 10_000), (2, 10_000), (2, 5_000)]  # two groups per call
    assert [r.repetitions for r in res] == [25_000] * 3
    assert res[0].mean == 1 and res[1].mean == -1
    s2 = OracleSampler(4)
    res2 = obs.measure_observables(c, [cirq.Z(q[1])], s2, obs.RepetitionsStoppingCriteria(1000),
                                   readout_symmetrization=True)
    assert s2.calls == [(2, 500)] and res2[0].mean == -1 and res2[0].repetitions == 1000


def test_variance_stopping_and_throttling():
    q = cirq.LineQubit.range(1)
    c = cirq.Circuit(cirq.ry(0.9).on(q[0]))
    s = OracleSampler(5)
    res = obs.measure_observables(c, [cirq.Z(q[0])], s, obs.VarianceStoppingCriteria(1e-4, repetitions_per_chunk=2000))
    assert res[0].variance <= 1e-4 and res[0].repetitions % 2000 == 0
    many = cirq.Points("t", list(np.linspace(0, 1, 400)))
    t = sympy.Symbol("t")
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        todo, reps = obs._still_todo(
            [obs._MeasurementSpec(None, ((i,),)) for i in range(400)],
            {obs._MeasurementSpec(None, ((i,),)): obs.BitstringAccumulator(None, [], {q[0]: 0}) for i in range(400)},
            obs.RepetitionsStoppingCriteria(100_000))
        assert reps == obs.MAX_REPETITIONS_PER_JOB // 400 and any("throttled" in str(x.message) for x in w)
    del many, t


def test_errors():
    q = cirq.LineQubit.range(1)
    c = cirq.Circuit(cirq.H(q[0]))
    with pytest.raises(ValueError):
        obs.sample_expectation_values(OracleSampler(), c, [cirq.Z(q[0])], num_samples=0)
    with pytest.raises(ValueError):
        obs.sample_expectation_values(OracleSampler(), c, [], num_samples=10)
    with pytest.raises(ValueError):
        obs.sample_expectation_values(OracleSampler(), cirq.Circuit(cirq.H(q[0]), cirq.measure(q[0])), [cirq.Z(q[0])],
                                      num_samples=10)
    with pytest.raises(ValueError):
        obs.sample_expectation_values(OracleSampler(), c, [1j * cirq.Z(q[0])], num_samples=10)
