"""Moment and Circuit containers (cirq-core/cirq/circuits/moment.py:51, circuit.py:1735).

Only what the simulators consume is mirrored: construction from op trees with Cirq's EARLIEST
insertion rule, iteration over moments and operations, qubit discovery, parameter queries,
``+``/``+=``/``*`` and resolution.
"""

from __future__ import annotations

import enum
from typing import Iterable, Iterator

from .gates import Operation, is_parameterized


def flatten_op_tree(tree) -> Iterator[Operation]:
    if isinstance(tree, Operation):
        yield tree
    elif isinstance(tree, Moment):
        yield from tree.operations
    elif isinstance(tree, Iterable) and not isinstance(tree, (str, bytes)):
        for sub in tree:
            yield from flatten_op_tree(sub)
    else:
        raise TypeError(f"Not an Operation or an iterable of operations: {tree!r}")


class InsertStrategy(enum.Enum):
    EARLIEST = 1
    NEW = 2
    INLINE = 3
    NEW_THEN_INLINE = 4


class Moment:
    """Operations that act on disjoint qubits during one time slice."""

    def __init__(self, *contents):
        ops = list(flatten_op_tree(contents))
        used = set()
        for op in ops:
            qs = set(op.qubits)
            if used & qs:
                raise ValueError(f"Overlapping operations: {ops}")
            used |= qs
        self._operations = tuple(ops)
        self._qubits = frozenset(used)

    @property
    def operations(self) -> tuple:
        return self._operations

    @property
    def qubits(self) -> frozenset:
        return self._qubits

    def operates_on(self, qubits) -> bool:
        return bool(self._qubits & set(qubits))

    def operation_at(self, qubit):
        for op in self._operations:
            if qubit in op.qubits:
                return op
        return None

    def with_operation(self, op) -> "Moment":
        return Moment(self._operations + (op,))

    def __iter__(self):
        return iter(self._operations)

    def __len__(self):
        return len(self._operations)

    def _is_parameterized_(self):
        return any(is_parameterized(op) for op in self._operations)

    def _resolve_parameters_(self, resolver, recursive=True):
        return Moment([_resolve(op, resolver, recursive) for op in self._operations])

    def __eq__(self, other):
        if not isinstance(other, Moment):
            return NotImplemented
        return set(self._operations) == set(other._operations)

    def __hash__(self):
        return hash(frozenset(self._operations))

    def __repr__(self):
        return f"cirq.Moment({', '.join(repr(o) for o in self._operations)})"


def _resolve(op, resolver, recursive=True):
    f = getattr(op, "_resolve_parameters_", None)
    if f is None or not is_parameterized(op):
        return op
    return f(resolver, recursive)


class Circuit:
    """A sequence of Moments."""

    def __init__(self, *contents, strategy: InsertStrategy = InsertStrategy.EARLIEST):
        self._moments: list[Moment] = []
        self.append(contents, strategy=strategy)

    @staticmethod
    def from_moments(*moments) -> "Circuit":
        c = Circuit()
        for m in moments:
            c._moments.append(m if isinstance(m, Moment) else Moment(m))
        return c

    @property
    def moments(self) -> tuple:
        return tuple(self._moments)

    def __iter__(self):
        return iter(self._moments)

    def __len__(self):
        return len(self._moments)

    def __getitem__(self, i):
        if isinstance(i, slice):
            return Circuit.from_moments(*self._moments[i])
        return self._moments[i]

    def all_operations(self) -> Iterator[Operation]:
        for m in self._moments:
            yield from m.operations

    def all_qubits(self) -> frozenset:
        qs = set()
        for m in self._moments:
            qs |= m.qubits
        return frozenset(qs)

    def _earliest_index(self, op, end: int) -> int:
        qs = set(op.qubits)
        k = end
        while k > 0 and not self._moments[k - 1].operates_on(qs):
            k -= 1
        return k

    def append(self, moment_or_operation_tree, strategy: InsertStrategy = InsertStrategy.EARLIEST):
        self.insert(len(self._moments), moment_or_operation_tree, strategy)

    def insert(self, index: int, tree, strategy: InsertStrategy = InsertStrategy.EARLIEST) -> int:
        """Inserts ops or moments (Moments are inserted whole), Cirq-style.  Returns the index after."""
        k = max(0, min(index, len(self._moments)))
        items = [tree] if isinstance(tree, (Operation, Moment)) else list(_moments_and_ops(tree))
        new_moment_created = False
        for it in items:
            if isinstance(it, Moment):
                self._moments.insert(k, it)
                k += 1
                new_moment_created = False
                continue
            op = it
            if strategy == InsertStrategy.EARLIEST:
                # earliest moment at or after the last moment touching op's qubits, before k
                j = self._earliest_index(op, k)
                while j < k and self._moments[j].operates_on(op.qubits):
                    j += 1
                if j < k:
                    self._moments[j] = self._moments[j].with_operation(op)
                    continue
                self._moments.insert(k, Moment([op]))
                k += 1
            elif strategy == InsertStrategy.NEW:
                self._moments.insert(k, Moment([op]))
                k += 1
            else:  # INLINE / NEW_THEN_INLINE
                if (
                    strategy == InsertStrategy.INLINE or new_moment_created
                ) and k > 0 and not self._moments[k - 1].operates_on(op.qubits):
                    self._moments[k - 1] = self._moments[k - 1].with_operation(op)
                else:
                    self._moments.insert(k, Moment([op]))
                    k += 1
                    new_moment_created = True
        return k

    def __iadd__(self, other):
        self.append(other)
        return self

    def __add__(self, other):
        c = self.copy()
        c.append(other if not isinstance(other, Circuit) else list(other.moments))
        return c

    def __radd__(self, other):
        c = Circuit(other)
        c.append(list(self._moments))
        return c

    def __mul__(self, reps: int):
        c = Circuit()
        for _ in range(int(reps)):
            c._moments.extend(self._moments)
        return c

    __rmul__ = __mul__

    def copy(self) -> "Circuit":
        return Circuit.from_moments(*self._moments)

    def _is_parameterized_(self):
        return any(m._is_parameterized_() for m in self._moments)

    def _parameter_names_(self) -> set:
        names = set()
        for op in self.all_operations():
            f = getattr(op, "_parameter_names_", None)
            if f is not None:
                names |= set(f())
        return names

    def _resolve_parameters_(self, resolver, recursive=True) -> "Circuit":
        return Circuit.from_moments(*[m._resolve_parameters_(resolver, recursive) for m in self._moments])

    def has_measurements(self) -> bool:
        return any(_is_measurement(op) for op in self.all_operations())

    def are_all_measurements_terminal(self) -> bool:
        for i, m in enumerate(self._moments):
            for op in m.operations:
                if _is_measurement(op):
                    later = self._moments[i + 1 :]
                    if any(mm.operates_on(op.qubits) for mm in later):
                        return False
        return True

    def are_any_measurements_terminal(self) -> bool:
        """True iff some measurement is followed by no operation on its qubits
        (cirq-core/cirq/circuits/circuit.py:889-894)."""
        for i, m in enumerate(self._moments):
            for op in m.operations:
                if _is_measurement(op):
                    later = self._moments[i + 1 :]
                    if not any(mm.operates_on(op.qubits) for mm in later):
                        return True
        return False

    def __eq__(self, other):
        if not isinstance(other, Circuit):
            return NotImplemented
        return self._moments == other._moments

    def __repr__(self):
        return f"cirq.Circuit({', '.join(repr(m) for m in self._moments)})"


def _is_measurement(op) -> bool:
    g = getattr(op, "gate", None)
    return g is not None and type(g).__name__ == "MeasurementGate"


def _moments_and_ops(tree):
    if isinstance(tree, (Operation, Moment)):
        yield tree
    elif isinstance(tree, Circuit):
        yield from tree.moments
    elif isinstance(tree, Iterable) and not isinstance(tree, (str, bytes)):
        for sub in tree:
            yield from _moments_and_ops(sub)
    else:
        raise TypeError(f"Not an Operation or an iterable of operations: {tree!r}")
