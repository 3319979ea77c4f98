"""GPU tests of the batched sweep (cqs_batch_* through the C ABI): every point's expectation values
and adjoint gradients against the oracle (small circuits) and against the single-state GPU path,
with and without HIP-graph replay, over two successive parameter sets."""

import sys

import numpy as np
import pytest

import cirq_amd as cirq
from cirq_amd import engine, generators
from cirq_amd.adjoint import AdjointGradient
from cirq_amd.batch import BatchedExpectation
from cirq_amd.paulis import pauli_masks
from tests.test_adjoint_host import _oracle_ev, _oracle_grad
from tests.test_batch_host import _workload

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_small_matches_oracle(gpu, use_graph):
    _log(f"batch small use_graph={use_graph}")
    circuits, observables, values = _workload()
    be = BatchedExpectation(circuits, observables, tile_bits=4, adjoint_tile_bits=3, coalesce_bits=1,
                            use_graph=use_graph)
    for rep in range(2):
        vals = [v + 0.37 * rep for v in values]
        evs, grads = be.run(vals)
        for c, circuit in enumerate(circuits):
            qubits = sorted(circuit.all_qubits())
            syms = be.symbols[c]
            psum = cirq.PauliSum.from_pauli_strings(observables[c])
            for b in range(vals[c].shape[0]):
                res = cirq.ParamResolver(dict(zip(syms, vals[c][b])))
                for t, term in enumerate(observables[c]):
                    assert abs(evs[c][b, t] - _oracle_ev(circuit, qubits, cirq.PauliSum.wrap(term), res)) < 1e-5
                np.testing.assert_allclose(grads[c][b], _oracle_grad(circuit, qubits, psum, syms, vals[c][b]),
                                           atol=1e-5)


def test_batch_mixed_sizes_match_single_state_path(gpu):
    _log("batch mixed sizes vs single-state path")
    circuits, observables, values = generators.mixed_sweep_workload(
        n_circuits=6, points=3, n_terms=8, qubit_range=(9, 15), layer_range=(2, 5), seed=3)
    be = BatchedExpectation(circuits, observables)
    for rep in range(2):
        vals = [v * (1.0 - 0.3 * rep) for v in values]
        evs, grads = be.run(vals)
        for c, circuit in enumerate(circuits):
            psum = cirq.PauliSum.from_pauli_strings(observables[c])
            ag = AdjointGradient(circuit, psum)
            qmap = {q: i for i, q in enumerate(ag.prog.qubits)}
            for b in range(vals[c].shape[0]):
                value, g = ag.value_and_grad(vals[c][b])
                np.testing.assert_allclose(grads[c][b], g, atol=1e-5)
                assert abs(evs[c][b].sum() - value) < 1e-5
                terms = [pauli_masks(t, qmap, ag.n) for t in observables[c]]
                single = engine.expect_terms(ag.forward_state(vals[c][b]), ag.n, False, terms, ag.ws).cpu().numpy()
                np.testing.assert_allclose(evs[c][b], single.real, atol=1e-6)
