"""Cirq-side plugin: the MI355X engine behind Cirq's own simulator interfaces.

In an installation where ``import cirq`` succeeds, ``AMDSimulator`` is a subclass of
``cirq.SimulatesSamples``, ``cirq.SimulatesFinalState`` and ``cirq.SimulatesExpectationValues``
(cirq-core/cirq/sim/simulator.py:97-116, 510-537, 398-440), as qsim registers itself
(docs/simulate/simulation.ipynb:661-676).  It takes real ``cirq.Circuit``, ``cirq.Sweepable``
and ``cirq.PauliSumLike`` objects, translates them through Cirq's JSON wire format
(``cirq.to_json`` -> ``cirq_amd.read_json``, the format pinned by the reference's own fixtures)
and runs them on ``cirq_amd.Simulator``:

    import cirq
    from cirq_amd.cirq_plugin import AMDSimulator
    result = AMDSimulator(seed=1234).run(cirq_circuit, repetitions=1000)   # a cirq.ResultDict

The translation helpers below need only JSON text and work without Cirq; they are what the
tests exercise here, where Cirq cannot be imported (Python 3.10).
"""

from __future__ import annotations

import numpy as np

from .json_io import read_json
from .simulator import Simulator
from .study import ParamResolver

try:  # the plugin classes exist only where Cirq does
    import cirq  # type: ignore
except Exception:  # ImportError, or a Cirq that fails to import on this interpreter
    cirq = None


def circuit_from_json(text: str):
    """A cirq_amd circuit from the JSON Cirq writes for a cirq.Circuit (or FrozenCircuit)."""
    return read_json(json_text=text)


def qubits_from_json(text: str) -> list:
    return list(read_json(json_text=text))


def resolver_from_json(text: str) -> ParamResolver:
    r = read_json(json_text=text)
    return r if isinstance(r, ParamResolver) else ParamResolver(r)


def observables_from_json(text: str) -> list:
    obs = read_json(json_text=text)
    return obs if isinstance(obs, list) else [obs]


def _to_json(obj) -> str:
    return cirq.to_json(obj)


def _resolvers(params) -> list:
    return [resolver_from_json(_to_json(r)) for r in cirq.to_resolvers(params)]


def _qubit_order(qubit_order, circuit) -> list:
    qs = cirq.QubitOrder.as_qubit_order(qubit_order).order_for(circuit.all_qubits())
    return qubits_from_json(_to_json(list(qs)))


if cirq is not None:

    class AMDSimulator(cirq.SimulatesSamples, cirq.SimulatesFinalState, cirq.SimulatesExpectationValues):
        """cirq_amd.Simulator (complex64 on one MI355X) as a Cirq simulator.  `noise` is a Cirq
        noise model applied on the Cirq side before translation."""

        def __init__(self, *, dtype=np.complex64, noise=None, seed=None, **engine_kwargs):
            self.noise = cirq.NoiseModel.from_noise_model_like(noise)
            self._engine = Simulator(dtype=dtype, seed=seed, **engine_kwargs)

        def _translate(self, circuit):
            if self.noise is not cirq.NO_NOISE:
                circuit = cirq.Circuit(self.noise.noisy_moments(circuit, sorted(circuit.all_qubits())))
            return circuit, circuit_from_json(_to_json(circuit))

        def _run(self, circuit, param_resolver, repetitions):
            _, c = self._translate(circuit)
            r = self._engine.run(c, resolver_from_json(_to_json(cirq.ParamResolver(param_resolver))), repetitions)
            return dict(r.records)

        def simulate_sweep_iter(self, program, params, qubit_order=cirq.QubitOrder.DEFAULT, initial_state=None):
            program, c = self._translate(program)
            qubits = cirq.QubitOrder.as_qubit_order(qubit_order).order_for(program.all_qubits())
            order = qubits_from_json(_to_json(list(qubits)))
            for r, res in zip(cirq.to_resolvers(params),
                              self._engine.simulate_sweep_iter(c, _resolvers(params), order, initial_state)):
                state = cirq.StateVectorSimulationState(
                    initial_state=res.final_state_vector, qubits=qubits, dtype=np.complex64)
                yield cirq.StateVectorTrialResult(params=r, measurements=dict(res.measurements),
                                                  final_simulator_state=state)

        def simulate_expectation_values_sweep_iter(self, program, observables, params,
                                                   qubit_order=cirq.QubitOrder.DEFAULT, initial_state=None,
                                                   permit_terminal_measurements=False):
            if not permit_terminal_measurements and program.are_any_measurements_terminal():
                raise ValueError("Provided circuit has terminal measurements, which may skew expectation values. "
                                 "If this is intentional, set permit_terminal_measurements=True.")
            program, c = self._translate(program)
            obs = observables if isinstance(observables, (list, tuple)) else [observables]
            ours = [observables_from_json(_to_json(cirq.PauliSum.wrap(o)))[0] for o in obs]
            yield from self._engine.simulate_expectation_values_sweep_iter(
                c, ours, _resolvers(params), _qubit_order(qubit_order, program), initial_state,
                permit_terminal_measurements=True)

    __all__ = ["AMDSimulator", "circuit_from_json", "observables_from_json", "qubits_from_json", "resolver_from_json"]
else:
    __all__ = ["circuit_from_json", "observables_from_json", "qubits_from_json", "resolver_from_json"]
