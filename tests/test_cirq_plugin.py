"""The Cirq-side plugin's translation layer, on JSON that Cirq itself wrote (the reference's own
serialization fixtures).  The plugin class needs an importable Cirq, which this interpreter does
not have (Cirq requires Python >= 3.11); its test runs only where Cirq imports."""

import json
import os

import numpy as np
import pytest

import cirq_amd
from cirq_amd import cirq_plugin as plugin

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "json")


def _text(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return f.read()


def test_circuits_translate_from_cirq_json():
    circuits = plugin.circuit_from_json(_text("Circuit.json"))
    assert [len(c) for c in circuits] == [2, 2, 1]
    assert isinstance(circuits[0], cirq_amd.Circuit)
    assert circuits[0].has_measurements() and len(circuits[0].all_qubits()) == 5


def test_resolvers_observables_and_qubits_translate():
    r = plugin.resolver_from_json(_text("ParamResolver.json"))
    assert isinstance(r, cirq_amd.ParamResolver) and r.value_of("a") == 0.1
    obs = plugin.observables_from_json(_text("PauliSum.json"))
    assert all(isinstance(o, cirq_amd.PauliSum) for o in obs)
    qs = plugin.qubits_from_json(_text("LineQubit.json")) + [plugin.circuit_from_json(_text("GridQubit.json"))]
    assert qs == [cirq_amd.LineQubit(0), cirq_amd.LineQubit(123), cirq_amd.GridQubit(10, 11)]


@pytest.mark.gpu
@pytest.mark.skipif(plugin.cirq is None, reason="Cirq does not import on this interpreter")
def test_plugin_simulator_matches_engine(gpu):  # pragma: no cover - needs Cirq
    cirq = plugin.cirq
    q = cirq.LineQubit.range(3)
    c = cirq.Circuit(cirq.H(q[0]), cirq.CNOT(q[0], q[1]), cirq.ry(0.3).on(q[2]), cirq.measure(*q, key="m"))
    sim = plugin.AMDSimulator(seed=5)
    assert isinstance(sim, cirq.SimulatesSamples)
    got = sim.run(c, repetitions=100).measurements["m"]
    ours = cirq_amd.Simulator(seed=5).run(plugin.circuit_from_json(cirq.to_json(c)), repetitions=100)
    assert np.array_equal(got, ours.measurements["m"])
