"""Qubit identifiers with Cirq's ordering and string forms.

Mirrors cirq-core/cirq/devices/line_qubit.py:278, cirq-core/cirq/devices/grid_qubit.py:359 and
cirq-core/cirq/ops/named_qubit.py.  Sorting follows Cirq's Qid comparison: first by class name,
then by the class's comparison key (cirq-core/cirq/ops/raw_types.py, ``_cmp_tuple``), which is
what QubitOrder.DEFAULT uses (cirq-core/cirq/ops/qubit_order.py:150).
"""

from __future__ import annotations

import functools
import re
from typing import Iterable


def _pad_digits(text: str) -> str:
    """Natural-sort key of cirq.NamedQubit: runs of digits padded to 8 characters."""
    return re.sub(r"\d+", lambda m: m.group(0).rjust(8, "0"), text)


@functools.total_ordering
class Qid:
    """Base class: a two-level quantum object with Cirq's ordering rules."""

    dimension = 2

    def _comparison_key(self):
        raise NotImplementedError

    def _cmp_tuple(self):
        return (type(self).__name__, self._comparison_key())

    def __lt__(self, other):
        if not isinstance(other, Qid):
            return NotImplemented
        return self._cmp_tuple() < other._cmp_tuple()

    def __eq__(self, other):
        if not isinstance(other, Qid):
            return NotImplemented
        return self._cmp_tuple() == other._cmp_tuple()

    def __hash__(self):
        return hash(self._cmp_tuple())


class LineQubit(Qid):
    """A qubit on a line, ``cirq.LineQubit(x)``."""

    def __init__(self, x: int):
        self._x = int(x)

    @property
    def x(self) -> int:
        return self._x

    def _comparison_key(self):
        return (self._x,)

    @staticmethod
    def range(*args) -> list["LineQubit"]:
        return [LineQubit(i) for i in range(*args)]

    def __str__(self):
        return f"q({self._x})"

    def __repr__(self):
        return f"cirq.LineQubit({self._x})"

    def __add__(self, other: int) -> "LineQubit":
        return LineQubit(self._x + other)


class GridQubit(Qid):
    """A qubit on a 2D grid, ``cirq.GridQubit(row, col)``."""

    def __init__(self, row: int, col: int):
        self._row = int(row)
        self._col = int(col)

    @property
    def row(self) -> int:
        return self._row

    @property
    def col(self) -> int:
        return self._col

    def _comparison_key(self):
        return (self._row, self._col)

    @staticmethod
    def rect(rows: int, cols: int, top: int = 0, left: int = 0) -> list["GridQubit"]:
        return [GridQubit(top + r, left + c) for r in range(rows) for c in range(cols)]

    @staticmethod
    def square(diameter: int, top: int = 0, left: int = 0) -> list["GridQubit"]:
        return GridQubit.rect(diameter, diameter, top, left)

    def is_adjacent(self, other: "GridQubit") -> bool:
        return abs(self._row - other._row) + abs(self._col - other._col) == 1

    def __str__(self):
        return f"q({self._row}, {self._col})"

    def __repr__(self):
        return f"cirq.GridQubit({self._row}, {self._col})"


class NamedQubit(Qid):
    """A qubit identified by name, ``cirq.NamedQubit(name)``."""

    def __init__(self, name: str):
        self._name = str(name)

    @property
    def name(self) -> str:
        return self._name

    def _comparison_key(self):
        return (_pad_digits(self._name),)

    @staticmethod
    def range(*args, prefix: str) -> list["NamedQubit"]:
        return [NamedQubit(f"{prefix}{i}") for i in range(*args)]

    def __str__(self):
        return self._name

    def __repr__(self):
        return f"cirq.NamedQubit({self._name!r})"


class QubitOrder:
    """Qubit ordering policies (cirq-core/cirq/ops/qubit_order.py)."""

    def __init__(self, explicit: Iterable | None = None, fallback: "QubitOrder | None" = None):
        self._explicit = None if explicit is None else list(explicit)
        self._fallback = fallback

    @staticmethod
    def explicit(fixed_qubits: Iterable, fallback: "QubitOrder | None" = None) -> "QubitOrder":
        fixed = list(fixed_qubits)
        if len(set(fixed)) != len(fixed):
            raise ValueError("Qubits appear in fixed_order twice.")
        return QubitOrder(fixed, fallback)

    def order_for(self, qubits: Iterable) -> tuple:
        qubits = list(qubits)
        if self._explicit is None:
            return tuple(sorted(set(qubits)))
        fixed = list(self._explicit)
        extra = set(qubits) - set(fixed)
        if extra:
            if self._fallback is None:
                raise ValueError(f"Unexpected extra qubits: {extra}.")
            fixed += list(self._fallback.order_for(extra))
        return tuple(fixed)

    @staticmethod
    def as_qubit_order(val) -> "QubitOrder":
        if isinstance(val, QubitOrder):
            return val
        if val is None:
            return QubitOrder.DEFAULT
        if isinstance(val, Iterable):
            return QubitOrder.explicit(val)
        raise ValueError(f"Don't know how to interpret <{val}> as a Basis.")


QubitOrder.DEFAULT = QubitOrder()
