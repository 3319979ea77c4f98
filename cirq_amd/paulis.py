"""Pauli gates, PauliString and PauliSum with Cirq's algebra.

Mirrors cirq-core/cirq/ops/pauli_gates.py, cirq-core/cirq/ops/pauli_string.py and
cirq-core/cirq/ops/linear_combinations.py:352 (PauliSum).  ``cirq.X(q)`` is itself a PauliString
(a one-qubit Pauli string operation) that can be placed in a circuit or combined with ``*``/``+``
into observables.
"""

from __future__ import annotations

import math

import numbers
from typing import Iterable, Mapping

import numpy as np
import sympy

from .gates import GateOperation, Operation, XPowGate, YPowGate, ZPowGate


class Pauli:
    """Mixin for the three Pauli gates; ``_index`` is 0 (X), 1 (Y) or 2 (Z)."""

    _index = -1
    _name = "?"

    def on(self, *qubits):
        if len(qubits) != 1:
            raise ValueError(f"{self._name} acts on exactly one qubit")
        return SingleQubitPauliStringGateOperation(self, qubits[0])

    def __call__(self, *qubits):
        return self.on(*qubits)

    def __pow__(self, power):
        base = (XPowGate, YPowGate, ZPowGate)[self._index]
        return self if power == 1 else base(exponent=power)

    def __eq__(self, other):
        return isinstance(other, Pauli) and other._index == self._index

    def __hash__(self):
        return hash(("Pauli", self._index))

    def __repr__(self):
        return f"cirq.{self._name}"


class _PauliX(Pauli, XPowGate):
    _index, _name = 0, "X"

    def __init__(self):
        XPowGate.__init__(self, exponent=1.0)


class _PauliY(Pauli, YPowGate):
    _index, _name = 1, "Y"

    def __init__(self):
        YPowGate.__init__(self, exponent=1.0)


class _PauliZ(Pauli, ZPowGate):
    _index, _name = 2, "Z"

    def __init__(self):
        ZPowGate.__init__(self, exponent=1.0)


X = _PauliX()
Y = _PauliY()
Z = _PauliZ()
_PAULIS = (X, Y, Z)


def _mul_paulis(a: Pauli, b: Pauli):
    """a*b = phase * c (c None for the identity)."""
    if a._index == b._index:
        return 1, None
    k = 3 - a._index - b._index
    phase = 1j if (b._index - a._index) % 3 == 1 else -1j
    return phase, _PAULIS[k]


class PauliString(Operation):
    """coefficient * prod_q P_q (cirq-core/cirq/ops/pauli_string.py)."""

    def __init__(self, *contents, qubit_pauli_map: Mapping | None = None, coefficient=1):
        m: dict = {}
        coef = coefficient
        for c in contents:
            if isinstance(c, PauliString):
                coef = coef * c.coefficient
                for q, p in c.items():
                    ph, r = _mul_paulis(m[q], p) if q in m else (1, p)
                    coef = coef * ph
                    if r is None:
                        m.pop(q, None)
                    else:
                        m[q] = r
            elif isinstance(c, numbers.Number) or isinstance(c, sympy.Basic):
                coef = coef * c
            elif isinstance(c, Mapping):
                for q, p in c.items():
                    m[q] = p
            else:
                raise TypeError(f"cannot build a PauliString from {c!r}")
        if qubit_pauli_map is not None:
            for q, p in qubit_pauli_map.items():
                if not isinstance(p, Pauli):
                    raise TypeError(f"{p!r} is not a Pauli")
                m[q] = p
        self._map = dict(sorted(m.items(), key=lambda kv: kv[0]))
        self._coefficient = coef if isinstance(coef, sympy.Basic) else complex(coef)

    # ---- mapping view
    @property
    def coefficient(self):
        return self._coefficient

    def items(self):
        return self._map.items()

    def keys(self):
        return self._map.keys()

    def values(self):
        return self._map.values()

    def __getitem__(self, q):
        return self._map[q]

    def get(self, q, default=None):
        return self._map.get(q, default)

    def __len__(self):
        return len(self._map)

    def __iter__(self):
        return iter(self._map)

    def __contains__(self, q):
        return q in self._map

    @property
    def qubits(self) -> tuple:
        return tuple(self._map.keys())

    @property
    def gate(self):
        return None

    def with_coefficient(self, c) -> "PauliString":
        return PauliString(qubit_pauli_map=self._map, coefficient=c)

    def _key(self):
        return frozenset((q, p._index) for q, p in self._map.items())

    # ---- algebra
    def __mul__(self, other):
        if isinstance(other, PauliString):
            return PauliString(self, other)
        if isinstance(other, PauliSum):
            return PauliSum.from_pauli_strings([self]) * other
        if isinstance(other, (numbers.Number, sympy.Basic)):
            return PauliString(qubit_pauli_map=self._map, coefficient=self._coefficient * other)
        return NotImplemented

    def __rmul__(self, other):
        if isinstance(other, (numbers.Number, sympy.Basic)):
            return PauliString(qubit_pauli_map=self._map, coefficient=self._coefficient * other)
        if isinstance(other, PauliString):
            return PauliString(other, self)
        return NotImplemented

    def __truediv__(self, other):
        if isinstance(other, numbers.Number):
            return self * (1 / other)
        return NotImplemented

    def __neg__(self):
        return self * -1

    def __pos__(self):
        return self

    def __add__(self, other):
        return PauliSum.wrap(self) + other

    def __radd__(self, other):
        return PauliSum.wrap(other) + self

    def __sub__(self, other):
        return PauliSum.wrap(self) - other

    def __rsub__(self, other):
        return PauliSum.wrap(other) - self

    def __pow__(self, power):
        if power == 1:
            return self
        return NotImplemented

    def __eq__(self, other):
        if not isinstance(other, PauliString):
            return NotImplemented
        return self._coefficient == other._coefficient and self._key() == other._key()

    def __hash__(self):
        return hash((self._coefficient, self._key()))

    def _is_parameterized_(self):
        return isinstance(self._coefficient, sympy.Basic) and bool(self._coefficient.free_symbols)

    def _has_unitary_(self):
        return not self._is_parameterized_() and abs(abs(complex(self._coefficient)) - 1) < 1e-9

    def _unitary_(self):
        if not self._has_unitary_():
            return NotImplemented
        u = np.array([[complex(self._coefficient)]])
        for p in self._map.values():
            u = np.kron(u, p._unitary_())
        return u

    def __repr__(self):
        body = "*".join(f"cirq.{p._name}({q!r})" for q, p in self._map.items()) or "cirq.PauliString()"
        return body if self._coefficient == 1 else f"({self._coefficient!r})*{body}"


class SingleQubitPauliStringGateOperation(PauliString, GateOperation):
    """``cirq.X(q)``: both a PauliString and a gate operation of the Pauli gate."""

    def __init__(self, pauli: Pauli, qubit):
        PauliString.__init__(self, qubit_pauli_map={qubit: pauli})
        self._gate = pauli
        self._qubits = (qubit,)

    @property
    def gate(self):
        return self._gate

    @property
    def qubits(self):
        return self._qubits

    def __pow__(self, power):
        if power == 1:
            return self
        return GateOperation(self._gate**power, self._qubits)

    __eq__ = PauliString.__eq__
    __hash__ = PauliString.__hash__


class PauliSum:
    """A linear combination of Pauli strings (cirq-core/cirq/ops/linear_combinations.py:352)."""

    def __init__(self, terms: Mapping | None = None):
        # key: frozenset of (qubit, pauli index) -> coefficient
        self._terms: dict = {}
        if terms:
            for k, v in terms.items():
                if v != 0:
                    self._terms[k] = v

    @staticmethod
    def from_pauli_strings(terms) -> "PauliSum":
        if isinstance(terms, PauliString):
            terms = [terms]
        s = PauliSum()
        for t in terms:
            s._add_term(t)
        return s

    @staticmethod
    def wrap(val) -> "PauliSum":
        if isinstance(val, PauliSum):
            return val
        if isinstance(val, PauliString):
            return PauliSum.from_pauli_strings([val])
        if isinstance(val, (numbers.Number, sympy.Basic)):
            return PauliSum.from_pauli_strings([PauliString(coefficient=val)])
        raise TypeError(f"cannot interpret {val!r} as a PauliSum")

    def _add_term(self, t: PauliString, sign=1):
        k = t._key()
        v = self._terms.get(k, 0) + sign * t.coefficient
        if v == 0:
            self._terms.pop(k, None)
        else:
            self._terms[k] = v

    def copy(self) -> "PauliSum":
        return PauliSum(dict(self._terms))

    def __iter__(self):
        for k, v in self._terms.items():
            yield PauliString(qubit_pauli_map={q: _PAULIS[i] for q, i in k}, coefficient=v)

    def __len__(self):
        return len(self._terms)

    @property
    def qubits(self) -> tuple:
        qs = set()
        for k in self._terms:
            qs |= {q for q, _ in k}
        return tuple(sorted(qs))

    def __add__(self, other):
        try:
            other = PauliSum.wrap(other)
        except TypeError:
            return NotImplemented
        r = self.copy()
        for t in other:
            r._add_term(t)
        return r

    __radd__ = __add__

    def __sub__(self, other):
        try:
            other = PauliSum.wrap(other)
        except TypeError:
            return NotImplemented
        r = self.copy()
        for t in other:
            r._add_term(t, -1)
        return r

    def __rsub__(self, other):
        return PauliSum.wrap(other) - self

    def __neg__(self):
        return self * -1

    def __mul__(self, other):
        if isinstance(other, (numbers.Number, sympy.Basic)):
            return PauliSum({k: v * other for k, v in self._terms.items()})
        if isinstance(other, (PauliString, PauliSum)):
            other = PauliSum.wrap(other)
            r = PauliSum()
            for a in self:
                for b in other:
                    r._add_term(a * b)
            return r
        return NotImplemented

    def __rmul__(self, other):
        if isinstance(other, (numbers.Number, sympy.Basic)):
            return self * other
        if isinstance(other, PauliString):
            return PauliSum.wrap(other) * self
        return NotImplemented

    def __truediv__(self, other):
        return self * (1 / other)

    def __eq__(self, other):
        if not isinstance(other, PauliSum):
            return NotImplemented
        return self._terms == other._terms

    def __repr__(self):
        return " + ".join(repr(t) for t in self) or "cirq.PauliSum()"


def pauli_sum_like(val) -> PauliSum:
    """Interprets a PauliSumLike (number, PauliString or PauliSum)."""
    return PauliSum.wrap(val)


def adjoint_scale(coefs) -> float:
    """Power of two >= sum |c| (and >= 1).  Adjoint sweeps start from phi = (H / s) psi (or
    vec(H) / s) and multiply their gradients by s: every op is linear in phi, the division by a
    power of two is exact, and phi's entries stay within the unit range that the split-FP16 MFMA
    path scales for (|amplitude| <= 1; a larger entry would overflow FP16 there)."""
    total = float(sum(abs(complex(c)) for c in coefs))
    return float(2.0 ** max(0, math.ceil(math.log2(total)))) if total > 1 else 1.0


def pauli_masks(term: PauliString, qubit_index: Mapping, n: int):
    """(xmask, zmask, coefficient * i^nY) of one term, with qubit axis i at flat bit n-1-i.

    With P[i ^ x, i] = i^nY (-1)^popcount(i & z) for the X/Y qubits in x and Z/Y qubits in z,
    <psi|P|psi> = i^nY sum_i (-1)^popcount(i & z) conj(psi[i ^ x]) psi[i].
    """
    x = z = 0
    ny = 0
    for q, p in term.items():
        bit = 1 << (n - 1 - qubit_index[q])
        if p._index in (0, 1):
            x |= bit
        if p._index in (1, 2):
            z |= bit
        if p._index == 1:
            ny += 1
    return x, z, complex(term.coefficient) * (1j**ny)
