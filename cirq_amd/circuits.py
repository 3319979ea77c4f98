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
    if g is not None and type(g).__name__ == "MeasurementGate":
        return True
    f = getattr(op, "_is_measurement_", None)
    return bool(f()) if f is not None else False


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


class FrozenCircuit(Circuit):
    """An immutable Circuit (cirq-core/cirq/circuits/frozen_circuit.py): what a CircuitOperation
    wraps.  Appending to it raises."""

    def __init__(self, *contents, strategy: InsertStrategy = InsertStrategy.EARLIEST):
        super().__init__(*contents, strategy=strategy)
        self._frozen = True

    @staticmethod
    def from_moments(*moments) -> "FrozenCircuit":
        c = FrozenCircuit()
        c._frozen = False
        for m in moments:
            c._moments.append(m if isinstance(m, Moment) else Moment(m))
        c._frozen = True
        return c

    def insert(self, index: int, tree, strategy: InsertStrategy = InsertStrategy.EARLIEST) -> int:
        if getattr(self, "_frozen", False):
            raise TypeError("FrozenCircuit is immutable")
        return super().insert(index, tree, strategy)

    def freeze(self) -> "FrozenCircuit":
        return self

    def unfreeze(self) -> Circuit:
        return Circuit.from_moments(*self._moments)

    def __hash__(self):
        return hash(tuple(self._moments))


def _freeze(self) -> FrozenCircuit:
    return FrozenCircuit.from_moments(*self._moments)


Circuit.freeze = _freeze


class CircuitOperation(Operation):
    """A sub-circuit used as one operation (cirq-core/cirq/circuits/circuit_operation.py:72-260):
    the wrapped FrozenCircuit with its qubits remapped (`qubit_map`), measurement keys renamed
    (`measurement_key_map`), parameters resolved by `param_resolver`, and repeated `repetitions`
    times (a negative count repeats the inverse).  With `use_repetition_ids` (the default when
    `repetition_ids` are given) every measurement key of repetition r becomes "<id_r>:<key>".
    The simulators apply it through its decomposition, as Cirq's act_on falls back to
    decompose for it."""

    def __init__(self, circuit, repetitions: int = 1, qubit_map=None, measurement_key_map=None, param_resolver=None,
                 repetition_ids=None, use_repetition_ids=None):
        if not isinstance(circuit, FrozenCircuit):
            raise TypeError(f"Expected circuit of type FrozenCircuit, got: {type(circuit)!r}")
        if isinstance(repetitions, float) and repetitions.is_integer():
            repetitions = int(repetitions)
        if not isinstance(repetitions, (int,)) or isinstance(repetitions, bool):
            raise TypeError(f"Only integer repetitions are allowed; got {repetitions!r}")
        self.circuit = circuit
        self.repetitions = int(repetitions)
        self.qubit_map = dict(qubit_map or {})
        self.measurement_key_map = dict(measurement_key_map or {})
        for v in self.measurement_key_map.values():
            if ":" in str(v):
                raise ValueError(f"Mapping to invalid key: {v}. ':' is not allowed.")
        from .study import ParamResolver

        self.param_resolver = param_resolver if isinstance(param_resolver, ParamResolver) else ParamResolver(
            param_resolver or {})
        if use_repetition_ids is None:
            use_repetition_ids = repetition_ids is not None
        self.use_repetition_ids = bool(use_repetition_ids)
        if repetition_ids is None and abs(self.repetitions) > 1:
            repetition_ids = [str(i) for i in range(abs(self.repetitions))]
        if repetition_ids is not None and len(repetition_ids) != abs(self.repetitions):
            raise ValueError(f"Expected repetition_ids to be a list of length {abs(self.repetitions)}, "
                             f"got: {repetition_ids}")
        self.repetition_ids = None if repetition_ids is None else list(repetition_ids)

    @property
    def qubits(self) -> tuple:
        from .qubits import QubitOrder

        inner = QubitOrder.DEFAULT.order_for(self.circuit.all_qubits())
        return tuple(self.qubit_map.get(q, q) for q in inner)

    def with_qubits(self, *new_qubits) -> "CircuitOperation":
        old = self.qubits
        if len(new_qubits) != len(old):
            raise ValueError(f"Expected {len(old)} new qubits, got {len(new_qubits)}")
        from .qubits import QubitOrder

        inner = QubitOrder.DEFAULT.order_for(self.circuit.all_qubits())
        return self._replace(qubit_map=dict(zip(inner, new_qubits)))

    def _replace(self, **kw) -> "CircuitOperation":
        args = dict(circuit=self.circuit, repetitions=self.repetitions, qubit_map=self.qubit_map,
                    measurement_key_map=self.measurement_key_map, param_resolver=self.param_resolver,
                    repetition_ids=self.repetition_ids, use_repetition_ids=self.use_repetition_ids)
        args.update(kw)
        return CircuitOperation(**args)

    def repeat(self, repetitions: int) -> "CircuitOperation":
        return self._replace(repetitions=self.repetitions * int(repetitions), repetition_ids=None)

    def _is_parameterized_(self) -> bool:
        return is_parameterized(self._resolved_circuit())

    def _parameter_names_(self) -> set:
        return self._resolved_circuit()._parameter_names_()

    def _resolve_parameters_(self, resolver, recursive=True) -> "CircuitOperation":
        """Symbols left free by `param_resolver` take their values from `resolver`."""
        from .study import ParamResolver

        inner = dict(self.param_resolver.param_dict)
        outer = resolver if isinstance(resolver, ParamResolver) else ParamResolver(resolver)
        merged = {}
        for name in self.circuit._parameter_names_():
            if name in inner:
                v = inner[name]
                merged[name] = outer.value_of(v) if is_parameterized(v) else v
            else:
                merged[name] = outer.value_of(name)
        for k, v in inner.items():
            merged.setdefault(k, v)
        return self._replace(param_resolver=ParamResolver(merged))

    def _resolved_circuit(self) -> Circuit:
        if not self.param_resolver.param_dict:
            return self.circuit
        return self.circuit._resolve_parameters_(self.param_resolver, True)

    def _has_unitary_(self) -> bool:
        return False

    def _is_measurement_(self) -> bool:
        return any(_is_measurement(op) or bool(getattr(op, "_is_measurement_", lambda: False)())
                   for op in self.circuit.all_operations())

    def _map_op(self, op, rep_id):
        from .gates import MeasurementGate

        qs = tuple(self.qubit_map.get(q, q) for q in op.qubits)
        g = getattr(op, "gate", None)
        if g is not None and isinstance(g, MeasurementGate):
            key = self.measurement_key_map.get(g.key, g.key)
            if rep_id is not None:
                key = f"{rep_id}:{key}"
            return MeasurementGate(g.num_qubits(), key, g.invert_mask).on(*qs)
        return op.with_qubits(*qs) if qs != tuple(op.qubits) else op

    def _inner_moments(self) -> list:
        """The resolved sub-circuit's moments, nested CircuitOperations unrolled in place."""
        out = []
        for m in self._resolved_circuit():
            plain = [op for op in m.operations if not isinstance(op, CircuitOperation)]
            nested = [op for op in m.operations if isinstance(op, CircuitOperation)]
            if plain or not nested:
                out.append(Moment(*plain))
            for op in nested:
                out.extend(op.mapped_circuit().moments)
        return out

    def mapped_circuit(self) -> Circuit:
        """The sub-circuit as it acts: repetitions unrolled, qubits and keys mapped, parameters
        resolved (cirq-core/cirq/circuits/circuit_operation.py:436-480)."""
        moments = self._inner_moments()
        if self.repetitions < 0:
            inv = []
            for m in reversed(moments):
                ops = []
                for op in m.operations:
                    p = op.__pow__(-1) if hasattr(op, "__pow__") else NotImplemented
                    if p is NotImplemented or p is None:
                        raise ValueError(f"Cannot invert {op!r} for negative repetitions")
                    ops.append(p)
                inv.append(Moment(*ops))
            moments = inv
        out = []
        for r in range(abs(self.repetitions)):
            rep_id = self.repetition_ids[r] if (self.use_repetition_ids and self.repetition_ids) else None
            for m in moments:
                out.append(Moment(*[self._map_op(op, rep_id) for op in m.operations]))
        return Circuit.from_moments(*out)

    def _decompose_(self):
        return list(self.mapped_circuit().all_operations())

    def __eq__(self, other):
        if not isinstance(other, CircuitOperation):
            return NotImplemented
        return (self.circuit == other.circuit and self.repetitions == other.repetitions
                and self.qubit_map == other.qubit_map and self.measurement_key_map == other.measurement_key_map
                and self.param_resolver.param_dict == other.param_resolver.param_dict
                and self.repetition_ids == other.repetition_ids and self.use_repetition_ids == other.use_repetition_ids)

    def __hash__(self):
        return hash((self.circuit, self.repetitions))

    def __repr__(self):
        return f"cirq.CircuitOperation({self.circuit!r}, repetitions={self.repetitions})"


MAX_DIRECT_QUBITS = 5  # wider gates that can decompose are applied through their decomposition


def _needs_decompose(op) -> bool:
    """Cirq's act_on falls back to decomposition when an op has no unitary, mixture, Kraus
    representation or measurement (cirq-core/cirq/sim/state_vector_simulation_state.py:376-399,
    cirq-core/cirq/sim/density_matrix_simulation_state.py:304-325), and apply_unitary tries the
    decomposition before the matrix for more than 4 axes
    (cirq-core/cirq/protocols/apply_unitary_protocol.py:375-389)."""
    if isinstance(op, CircuitOperation):
        return True
    if _is_measurement(op):
        return False
    g = getattr(op, "gate", None)
    obj = g if g is not None else op
    dec = getattr(obj, "_decompose_", None)
    if dec is None:
        return False
    if type(obj).__name__ in ("PauliString", "SingleQubitPauliStringGateOperation"):
        return False
    has_u = getattr(obj, "_has_unitary_", None)
    unitary = bool(has_u()) if has_u is not None else (getattr(obj, "_unitary_", None) is not None)
    if unitary:
        return len(op.qubits) > MAX_DIRECT_QUBITS
    if is_parameterized(obj):
        return False
    return not (hasattr(obj, "_kraus_") or hasattr(obj, "_mixture_"))


def _decomposed(op) -> list:
    if isinstance(op, CircuitOperation):
        res = op._decompose_()
    else:
        g = getattr(op, "gate", None)
        res = g._decompose_(op.qubits) if g is not None else op._decompose_()
    if res is NotImplemented or res is None:
        return [op]
    out = []
    for sub in flatten_op_tree(res):
        out.extend(_decomposed(sub) if _needs_decompose(sub) else [sub])
    return out


def expand_for_simulation(circuit: Circuit) -> Circuit:
    """The circuit with every CircuitOperation and every op that only decomposes replaced by its
    (recursive) decomposition.  A moment holding such ops becomes the moment of its other ops
    followed by the decomposed ops, which act on disjoint qubits.  Returns `circuit` itself when
    nothing needs expanding."""
    if not any(_needs_decompose(op) for op in circuit.all_operations()):
        return circuit
    moments = []
    for m in circuit:
        plain = [op for op in m.operations if not _needs_decompose(op)]
        wide = [op for op in m.operations if _needs_decompose(op)]
        if plain or not wide:
            moments.append(Moment(*plain))
        for op in wide:
            sub = Circuit(_decomposed(op))
            moments.extend(sub.moments)
    return Circuit.from_moments(*moments)
