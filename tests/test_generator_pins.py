"""Pins of the circuit generators against the reference tests' seeded literals.

* experiments/random_quantum_circuit_generation_test.py:42-77: the seed-52 two-qubit circuit.  It
  goes through the same single-qubit layer factory (prng.randint per qubit, redraw while equal to
  the qubit's previous gate) as the grid circuits of the benchmark, so this fixes that factory's
  draw order and gate table.
* experiments/random_quantum_circuit_generation_test.py:207-232: the pairs that the A, B, C, D
  layers of HALF_GRID_STAGGERED_PATTERN activate on a 3x3 grid, in order.
"""

import numpy as np

import cirq_amd as cirq
from cirq_amd import generators as G


def _names(moment):
    out = {}
    for op in moment:
        g = op.gate
        if isinstance(g, cirq.PhasedXPowGate):
            out[op.qubits[0]] = f"PhX({g.phase_exponent})^{g.exponent}"
        elif isinstance(g, cirq.XPowGate):
            out[op.qubits[0]] = f"X^{g.exponent}"
        elif isinstance(g, cirq.YPowGate):
            out[op.qubits[0]] = f"Y^{g.exponent}"
        elif isinstance(g, cirq.CZPowGate):
            out[op.qubits] = f"CZ^{g.exponent}"
    return out


def test_two_qubit_circuit_seed_52_known_answer():
    q0, q1 = cirq.LineQubit.range(2)
    c = G.random_rotations_between_two_qubit_circuit(q0, q1, 4, seed=52)
    assert len(c) == 4 * 2 + 1 and c.all_qubits() == {q0, q1}
    c = G.random_rotations_between_two_qubit_circuit(q0, q1, 4, seed=52, add_final_single_qubit_layer=False)
    assert len(c) == 4 * 2
    want = [
        ("Y^0.5", "X^0.5"),
        ("PhX(0.25)^0.5", "Y^0.5"),
        ("Y^0.5", "X^0.5"),
        ("X^0.5", "PhX(0.25)^0.5"),
    ]
    for cycle, (a, b) in enumerate(want):
        one = _names(c[2 * cycle])
        assert (one[q0], one[q1]) == (a, b), cycle
        assert _names(c[2 * cycle + 1]) == {(q0, q1): "CZ^1.0"}


def test_half_grid_staggered_layers_on_3x3_known_answer():
    gq = cirq.GridQubit
    qs = gq.rect(3, 3)
    want = [
        [(gq(0, 0), gq(1, 0)), (gq(0, 2), gq(1, 2)), (gq(1, 1), gq(2, 1))],  # A: vertical
        [(gq(0, 1), gq(1, 1)), (gq(1, 0), gq(2, 0)), (gq(1, 2), gq(2, 2))],  # B: vertical, offset
        [(gq(0, 1), gq(0, 2)), (gq(1, 0), gq(1, 1)), (gq(2, 1), gq(2, 2))],  # C: horizontal, offset
        [(gq(0, 0), gq(0, 1)), (gq(1, 1), gq(1, 2)), (gq(2, 0), gq(2, 1))],  # D: horizontal
    ]
    for layer, pairs in zip(G.HALF_GRID_STAGGERED_PATTERN, want):
        assert G.grid_interaction_layer_pairs(qs, layer) == sorted(pairs)


def test_grid_circuit_draws_follow_the_shared_factory():
    """The grid circuit's single-qubit layers are the factory's layers on the row-major qubits,
    interleaved with the pattern's CZ layers (same prng stream, CZ draws nothing)."""
    qs = cirq.GridQubit.rect(3, 3)
    c = G.grid_circuit(3, 3, 6, seed=7)
    new_layer = G._single_qubit_layer_factory((G.SQRT_X, G.SQRT_Y, G.SQRT_W), qs, np.random.RandomState(7))
    prev = cirq.Moment()
    moments = list(c)
    k = 0
    for cycle in range(6):
        layer = new_layer(prev)
        assert _names(moments[k]) == _names(layer)
        pairs = G.grid_interaction_layer_pairs(qs, G.HALF_GRID_STAGGERED_PATTERN[cycle % 4])
        assert sorted(_names(moments[k + 1])) == sorted(pairs)
        k += 2
        prev = layer
