"""Parameter expressions that are not affine in one symbol (several symbols, products, sympy
functions) resolve to the values the expressions fix, and a circuit built from them simulates to the
state of the same circuit with those numbers written in (CPU pass model against the oracle).

Cirq resolves such expressions through its ParamResolver (cirq-core/cirq/study/resolver.py:120-200):
the substituted sympy expression evaluated to a float.  The engine evaluates them with
sympy.lambdify over the expression's own symbols; both are the expression's value, which is what
these tests pin (closed forms computed with the math module)."""

import math

import numpy as np
import sympy

import cirq_amd as cirq
from cirq_amd import tape
from oracle import cirq_np as O
from tests.helpers import fidelity, run_pass_model, to_oracle_ops

a, b = sympy.symbols("a b")
EXPRS = [
    (a * b, lambda x, y: x * y),
    (sympy.sin(a) * b + a**2, lambda x, y: math.sin(x) * y + x * x),
    (sympy.exp(-a) / (1 + b**2), lambda x, y: math.exp(-x) / (1 + y * y)),
    (sympy.pi * a * b / 2, lambda x, y: math.pi * x * y / 2),
    (2 * a + 3 * b - 1, lambda x, y: 2 * x + 3 * y - 1),  # affine, but in two symbols
    (sympy.cos(a + b) ** 2, lambda x, y: math.cos(x + y) ** 2),
]
POINTS = [(0.3, -1.1), (2.5, 0.7), (-1.9, 2.2), (0.0, 0.0)]


def _evaluator(circuit, qubits):
    low = tape.Lowering(qubits, False)
    low.lower_circuit(circuit)
    return low, tape.ParamEvaluator(low.param_refs)


def _symvals(ev, points):
    return np.array([[{"a": x, "b": y}[s.name] for x, y in points] for s in ev.symbols], dtype=np.float64)


def test_evaluator_matches_closed_forms():
    circ, q = _circuit(e for e, _ in EXPRS)
    _, ev = _evaluator(circ, q)
    assert [s.name for s in ev.symbols] == ["a", "b"]
    symvals = _symvals(ev, POINTS)  # all points at once (one column per resolver)
    for e, f in EXPRS:
        got = ev.values(e, symvals)
        want = np.array([f(x, y) for x, y in POINTS])
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)


def _circuit(exprs):
    q = cirq.LineQubit.range(4)
    e = list(exprs)
    return cirq.Circuit(
        [cirq.H(q[i]) for i in range(4)]
        + [cirq.rx(e[0]).on(q[0]), cirq.rz(e[1]).on(q[1]), cirq.ZPowGate(exponent=e[2]).on(q[2]),
           cirq.CZ(q[0], q[1]), cirq.XXPowGate(exponent=e[3]).on(q[2], q[3]),
           cirq.ry(e[4]).on(q[3]), cirq.CZPowGate(exponent=e[5]).on(q[1], q[2]),
           cirq.FSimGate(theta=e[0], phi=e[5]).on(q[0], q[3])]
    ), q


def test_nonaffine_circuit_matches_the_numeric_circuit():
    sym_circ, q = _circuit(e for e, _ in EXPRS)
    low, ev = _evaluator(sym_circ, q)
    packed = tape.compile_ops(low.ops, 4, L=3, coalesce=1)
    mats = tape.fill_param_mats(packed, ev, _symvals(ev, POINTS))
    for i, (x, y) in enumerate(POINTS):
        st = np.zeros(16, dtype=np.complex128)
        st[0] = 1
        got = run_pass_model(packed, st, mats[i])
        num_circ, _ = _circuit(f(x, y) for _, f in EXPRS)
        want = O.simulate_state_vector(to_oracle_ops(num_circ, q), 4)
        assert np.max(np.abs(got - want)) <= 1e-6
        assert fidelity(got, want) >= 1 - 1e-9
