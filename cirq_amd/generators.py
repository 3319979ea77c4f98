"""Circuit generators for the benchmark workloads.

random_rotations_between_grid_interaction_layers_circuit restates
cirq-core/cirq/experiments/random_quantum_circuit_generation.py:538-699 (GridInteractionLayer at
:38-117, HALF_GRID_STAGGERED_PATTERN at :139-144, the single-qubit layer factory at :640-660 that
never repeats a qubit's previous gate by identity, coupled pairs at :625-637).  The QAOA gate
convention follows examples/qaoa.py:128-146.
"""

from __future__ import annotations

import dataclasses
from typing import Callable, Sequence

import numpy as np
import sympy

from . import gates as g
from .channels import amplitude_damp, depolarize
from .circuits import Circuit, Moment
from .paulis import X as PX
from .paulis import Y as PY
from .paulis import Z as PZ
from .paulis import PauliString, PauliSum
from .qubits import GridQubit, LineQubit


@dataclasses.dataclass(frozen=True)
class GridInteractionLayer:
    """A layer of aligned or staggered two-qubit interactions on a grid."""

    col_offset: int = 0
    vertical: bool = False
    stagger: bool = False

    def __contains__(self, pair) -> bool:
        a, b = pair
        if self.vertical:
            a, b = GridQubit(a.col, a.row), GridQubit(b.col, b.row)
        a, b = sorted((a, b))
        if a.row != b.row or b.col != a.col + 1:
            return False
        pos = (a.row % 2, (a.col - self.col_offset) % 2)
        return pos == (0, 0) or pos == (1, int(self.stagger))


GRID_STAGGERED_PATTERN = (
    GridInteractionLayer(col_offset=0, vertical=True, stagger=True),  # A
    GridInteractionLayer(col_offset=1, vertical=True, stagger=True),  # B
    GridInteractionLayer(col_offset=1, vertical=False, stagger=True),  # C
    GridInteractionLayer(col_offset=0, vertical=False, stagger=True),  # D
    GridInteractionLayer(col_offset=1, vertical=False, stagger=True),  # C
    GridInteractionLayer(col_offset=0, vertical=False, stagger=True),  # D
    GridInteractionLayer(col_offset=0, vertical=True, stagger=True),  # A
    GridInteractionLayer(col_offset=1, vertical=True, stagger=True),  # B
)

HALF_GRID_STAGGERED_PATTERN = GRID_STAGGERED_PATTERN[:4]


def _coupled_qubit_pairs(qubits: list) -> list:
    pairs = []
    qs = set(qubits)
    for q in qubits:
        for nb in (GridQubit(q.row, q.col + 1), GridQubit(q.row + 1, q.col)):
            if nb in qs:
                pairs.append((q, nb))
    return pairs


# The default gate set of the generator: X**0.5, Y**0.5 and PhasedX(0.25, 0.5) (sqrt-W).
SQRT_X = g.XPowGate(exponent=0.5)
SQRT_Y = g.YPowGate(exponent=0.5)
SQRT_W = g.PhasedXPowGate(phase_exponent=0.25, exponent=0.5)


def _single_qubit_layer_factory(single_qubit_gates: Sequence, qubits: Sequence, prng: np.random.RandomState):
    """_single_qubit_gates_arg_to_factory (random_quantum_circuit_generation.py:640-687): one
    distinct gate (by equality) is used everywhere; otherwise each layer draws a gate per qubit
    with prng.randint, redrawing while it is the qubit's previous gate (by identity)."""
    try:
        fixed = len(set(single_qubit_gates)) == 1
    except TypeError:  # unhashable gates: compare by identity
        fixed = len(set(map(id, single_qubit_gates))) == 1

    def new_layer(prev: Moment) -> Moment:
        if fixed:
            return Moment(single_qubit_gates[0].on(q) for q in qubits)
        ops = []
        for q in qubits:
            ex = prev.operation_at(q)
            excluded = ex.gate if ex is not None else None
            gate = single_qubit_gates[prng.randint(0, len(single_qubit_gates))]
            while gate is excluded:
                gate = single_qubit_gates[prng.randint(0, len(single_qubit_gates))]
            ops.append(gate.on(q))
        return Moment(ops)

    return new_layer


def _prng(seed) -> np.random.RandomState:
    return seed if isinstance(seed, np.random.RandomState) else np.random.RandomState(seed)


def random_rotations_between_grid_interaction_layers_circuit(
    qubits: Sequence[GridQubit],
    depth: int,
    *,
    two_qubit_op_factory: Callable = lambda a, b, _: g.CZPowGate().on(a, b),
    pattern: Sequence[GridInteractionLayer] = GRID_STAGGERED_PATTERN,
    single_qubit_gates: Sequence = (SQRT_X, SQRT_Y, SQRT_W),
    add_final_single_qubit_layer: bool = True,
    seed=None,
) -> Circuit:
    prng = _prng(seed)
    qubits = list(qubits)
    pairs = _coupled_qubit_pairs(qubits)
    circuit = Circuit()
    previous = Moment()
    new_layer = _single_qubit_layer_factory(single_qubit_gates, qubits, prng)
    for i in range(depth):
        layer = new_layer(previous)
        circuit += layer
        layer2 = pattern[i % len(pattern)]
        circuit += [two_qubit_op_factory(a, b, prng) for a, b in pairs if (a, b) in layer2 or (b, a) in layer2]
        previous = layer
    if add_final_single_qubit_layer:
        circuit += new_layer(previous)
    return circuit


def random_rotations_between_two_qubit_circuit(
    q0,
    q1,
    depth: int,
    two_qubit_op_factory: Callable = lambda a, b, _: g.CZPowGate().on(a, b),
    single_qubit_gates: Sequence = (SQRT_X, SQRT_Y, SQRT_W),
    add_final_single_qubit_layer: bool = True,
    seed=None,
) -> Circuit:
    """random_rotations_between_two_qubit_circuit (random_quantum_circuit_generation.py:171-231):
    `depth` cycles of a random single-qubit layer on (q0, q1) then the two-qubit op, with the
    same single-qubit layer factory (and draw order) as the grid generator."""
    prng = _prng(seed)
    circuit = Circuit()
    previous = Moment()
    new_layer = _single_qubit_layer_factory(single_qubit_gates, (q0, q1), prng)
    for _ in range(depth):
        layer = new_layer(previous)
        circuit += layer
        circuit += two_qubit_op_factory(q0, q1, prng)
        previous = layer
    if add_final_single_qubit_layer:
        circuit += new_layer(previous)
    return circuit


def grid_interaction_layer_pairs(qubits: Sequence[GridQubit], layer: GridInteractionLayer) -> list:
    """The nearest-neighbour pairs of `qubits` that `layer` activates, sorted (the pairs of one
    moment of get_grid_interaction_layer_circuit, random_quantum_circuit_generation.py:511-535)."""
    return sorted((a, b) for a, b in _coupled_qubit_pairs(list(qubits)) if (a, b) in layer or (b, a) in layer)


def grid_circuit(rows: int, cols: int, depth: int, seed: int = 0, n_qubits: int | None = None) -> Circuit:
    """The benchmark random circuit on a rows x cols grid (optionally its first n_qubits, row-major)."""
    qs = GridQubit.rect(rows, cols)
    if n_qubits is not None:
        qs = qs[:n_qubits]
    return random_rotations_between_grid_interaction_layers_circuit(
        qs, depth, pattern=HALF_GRID_STAGGERED_PATTERN, add_final_single_qubit_layer=False, seed=seed
    )


def nearest_neighbour_zz(qubits: Sequence[GridQubit]) -> list:
    """<Z_i Z_j> observables for every nearest-neighbour pair of a grid."""
    return [PZ(a) * PZ(b) for a, b in _coupled_qubit_pairs(list(qubits))]


# ---- noisy density-matrix workload ------------------------------------------------------------
def noisy_line_circuit(n: int, layers: int, seed: int = 0, p: float = 1e-3, gamma: float = 1e-3) -> Circuit:
    """Rz(theta) Ry(phi) per qubit, CZ on even/odd pairs alternating; depolarize(p) after every gate
    (on each qubit it touches) and amplitude_damp(gamma) on every qubit each layer."""
    rng = np.random.RandomState(seed)
    qs = LineQubit.range(n)
    c = Circuit()
    for layer in range(layers):
        th = rng.uniform(0, 2 * np.pi, n)
        ph = rng.uniform(0, 2 * np.pi, n)
        c.append(Moment(g.rz(th[i]).on(q) for i, q in enumerate(qs)))
        c.append(Moment(depolarize(p).on(q) for q in qs))
        c.append(Moment(g.ry(ph[i]).on(q) for i, q in enumerate(qs)))
        c.append(Moment(depolarize(p).on(q) for q in qs))
        start = layer % 2
        pairs = [(qs[i], qs[i + 1]) for i in range(start, n - 1, 2)]
        if pairs:
            c.append(Moment(g.CZ.on(a, b) for a, b in pairs))
            c.append(Moment(depolarize(p).on(q) for ab in pairs for q in ab))
        c.append(Moment(amplitude_damp(gamma).on(q) for q in qs))
    return c


def noisy_line_ansatz(n: int, layers: int, seed: int = 0, p: float = 1e-3, gamma: float = 1e-3):
    """noisy_line_circuit with every Rz / Ry angle a symbol.  Returns (circuit, symbols, values),
    the values drawn as noisy_line_circuit draws its angles."""
    rng = np.random.RandomState(seed)
    qs = LineQubit.range(n)
    c = Circuit()
    syms, vals = [], []
    for layer in range(layers):
        th = [sympy.Symbol(f"th{layer}_{i}") for i in range(n)]
        ph = [sympy.Symbol(f"ph{layer}_{i}") for i in range(n)]
        vals += list(rng.uniform(0, 2 * np.pi, n)) + list(rng.uniform(0, 2 * np.pi, n))
        syms += th + ph
        c.append(Moment(g.rz(th[i]).on(q) for i, q in enumerate(qs)))
        c.append(Moment(depolarize(p).on(q) for q in qs))
        c.append(Moment(g.ry(ph[i]).on(q) for i, q in enumerate(qs)))
        c.append(Moment(depolarize(p).on(q) for q in qs))
        start = layer % 2
        pairs = [(qs[i], qs[i + 1]) for i in range(start, n - 1, 2)]
        if pairs:
            c.append(Moment(g.CZ.on(a, b) for a, b in pairs))
            c.append(Moment(depolarize(p).on(q) for ab in pairs for q in ab))
        c.append(Moment(amplitude_damp(gamma).on(q) for q in qs))
    return c, syms, np.array(vals)


def sum_z(qubits) -> PauliSum:
    return PauliSum.from_pauli_strings([PZ(q) for q in qubits])


# ---- variational workloads --------------------------------------------------------------------
def hardware_efficient_ansatz(n: int, layers: int, seed: int = 0, prefix: str = "t"):
    """Per layer Ry(a_i) Rz(b_i) on every qubit then a CZ ladder.  Returns (circuit, symbols, values)."""
    qs = LineQubit.range(n)
    syms = []
    c = Circuit()
    for layer in range(layers):
        a = [sympy.Symbol(f"{prefix}{layer}_y{i}") for i in range(n)]
        b = [sympy.Symbol(f"{prefix}{layer}_z{i}") for i in range(n)]
        syms += a + b
        c.append(Moment(g.ry(a[i]).on(q) for i, q in enumerate(qs)))
        c.append(Moment(g.rz(b[i]).on(q) for i, q in enumerate(qs)))
        for i in range(0, n - 1, 2):
            c.append(g.CZ.on(qs[i], qs[i + 1]))
        for i in range(1, n - 1, 2):
            c.append(g.CZ.on(qs[i], qs[i + 1]))
    rng = np.random.RandomState(seed)
    vals = rng.uniform(-np.pi, np.pi, len(syms))
    return c, syms, vals


def random_pauli_sum(qubits, n_terms: int, locality: int, seed: int = 0, include_z: bool = True) -> PauliSum:
    """sum_i Z_i (optional) + n_terms random `locality`-local strings with N(0, 1) weights."""
    rng = np.random.RandomState(seed)
    qubits = list(qubits)
    terms = [PZ(q) for q in qubits] if include_z else []
    paulis = (PX, PY, PZ)
    for _ in range(n_terms):
        sup = rng.choice(len(qubits), size=locality, replace=False)
        w = rng.normal()
        m = {qubits[i]: paulis[rng.randint(3)] for i in sup}
        terms.append(PauliString(qubit_pauli_map=m, coefficient=w))
    return PauliSum.from_pauli_strings(terms)


def mixed_sweep_workload(n_circuits: int = 256, points: int = 16, n_terms: int = 32, locality: int = 3,
                         qubit_range=(10, 22), layer_range=(4, 16), seed: int = 0):
    """A mixed-size sweep: n_circuits hardware-efficient ansatz circuits whose qubit and layer counts
    are uniform in the given inclusive ranges, `points` rows of symbol values ~ U[-pi, pi] each, and
    per circuit n_terms random `locality`-local Pauli strings with N(0, 1) weights.
    Returns (circuits, observables: lists of PauliStrings, values: (points, n_symbols) arrays)."""
    rng = np.random.RandomState(seed)
    circuits, observables, values = [], [], []
    for k in range(n_circuits):
        n = int(rng.randint(qubit_range[0], qubit_range[1] + 1))
        layers = int(rng.randint(layer_range[0], layer_range[1] + 1))
        c, syms, _ = hardware_efficient_ansatz(n, layers, seed=k, prefix=f"c{k}_")
        qs = sorted(c.all_qubits())
        terms = list(random_pauli_sum(qs, n_terms, locality, seed=seed * 100003 + k, include_z=False))
        circuits.append(c)
        observables.append(terms)
        values.append(rng.uniform(-np.pi, np.pi, (points, len(syms))))
    return circuits, observables, values


def random_3_regular_graph(n: int, seed: int = 0) -> list:
    """Edges of a random 3-regular graph on n nodes (configuration model with rejection)."""
    rng = np.random.RandomState(seed)
    while True:
        stubs = np.repeat(np.arange(n), 3)
        rng.shuffle(stubs)
        edges = set()
        ok = True
        for a, b in zip(stubs[0::2], stubs[1::2]):
            a, b = int(min(a, b)), int(max(a, b))
            if a == b or (a, b) in edges:
                ok = False
                break
            edges.add((a, b))
        if ok:
            return sorted(edges)


def qaoa_maxcut(n: int, p: int, seed: int = 0):
    """QAOA for MaxCut on a random 3-regular graph: H on all, then per layer rzz(-gamma/2)-style
    ZZPowGate(2*(-0.5*gamma)/pi, global_shift=-1/2) on edges and rx(2*beta) on nodes
    (examples/qaoa.py:128-146).  Returns (circuit, symbols, values, cost PauliSum, edges)."""
    edges = random_3_regular_graph(n, seed)
    qs = LineQubit.range(n)
    gam = [sympy.Symbol(f"gamma{k}") for k in range(p)]
    bet = [sympy.Symbol(f"beta{k}") for k in range(p)]
    c = Circuit(g.H.on_each(*qs))
    for k in range(p):
        for a, b in edges:
            c.append(g.ZZPowGate(exponent=2 * (-0.5 * gam[k]) / sympy.pi, global_shift=-0.5).on(qs[a], qs[b]))
        c.append(Moment(g.rx(2 * bet[k]).on(q) for q in qs))
    rng = np.random.RandomState(seed + 1)
    vals = rng.uniform(0, np.pi, 2 * p)
    syms = gam + bet
    cost = PauliSum()
    for a, b in edges:
        cost = cost + 0.5 - 0.5 * PZ(qs[a]) * PZ(qs[b])
    return c, syms, vals, cost, edges
