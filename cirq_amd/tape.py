"""Circuit -> op tape -> passes: the host-side compiler of the engine.

Lowering replaces the reference's per-operation dispatch (cirq-core/cirq/sim/simulator_base.py:
198-212 -> protocols.act_on -> apply_unitary / apply_channel).  Instead of one NumPy pass per
operation, operations become bit-level ops (dense matrices on flat-index bits, diagonals, pair
phases, global scales) that the scheduler packs into passes: each pass reads and writes the state
once and applies every op whose non-diagonal targets fit in that pass's tile of L local bits.

Bit conventions: a state vector over qubit_order has axis i at flat bit n-1-i; a density matrix
has row axis i at bit 2n-1-i and column axis i at bit n-1-i.  A Cirq matrix on qubits
(q_0..q_{k-1}) is big-endian, so its kernel target list is (bit(q_{k-1}), ..., bit(q_0)).
"""

from __future__ import annotations

import dataclasses
from typing import Callable, Sequence

import numpy as np
import sympy

from . import _lib
from .gates import is_parameterized

MAX_DENSE = 5
MAX_DIAG = 6
OUTER = 64  # DIAG / PAIRPHASE bit encoding: tile-local index, or OUTER + global position


@dataclasses.dataclass
class ParamRef:
    """How to compute one lowered op's matrix from resolved parameter values.

    gate: the (unresolved) gate; transform: maps the gate's (B, d, d) matrix batch to this op's
    payload batch; exprs: the gate's parameter expressions; slot: first gradient slot.
    """

    gate: object
    transform: Callable
    exprs: tuple
    slot: int = -1
    deriv: int | None = None  # evaluate d(matrix)/d(parameter `deriv`) instead of the matrix


@dataclasses.dataclass
class LOp:
    kind: str  # 'dense' | 'diag' | 'pair' | 'scale'
    targets: tuple  # flat bit positions; matrix index bit m <-> targets[m]
    mat: np.ndarray | None  # dense: (d, d); diag: (d,); pair: scalar phase; scale: scalar
    controls: tuple = ()  # ((bit, value), ...) for dense ops
    pref: ParamRef | None = None
    sel: int = 0  # adjoint programs: 0 both states, 1 psi only, 2 phi only
    slot: int = -1  # adjoint GRAD ops: gradient slot
    vbits: tuple = ()  # fused dense / block ops: outer bits selecting one of 2^len matrices (mat[v])
    fence: bool = False  # scheduling barrier: runs after every earlier op, before every later one
    need: tuple = ()  # bits that must be local in the op's pass besides its targets

    @property
    def bits(self) -> frozenset:
        return frozenset(self.targets) | frozenset(b for b, _ in self.controls)

    @property
    def fixed(self) -> bool:
        return self.pref is None and self.sel == 0 and not self.kind.startswith("grad")


@dataclasses.dataclass
class Measurement:
    key: str
    qubits: tuple
    invert_mask: tuple


class UnsupportedOperation(TypeError):
    pass


# --------------------------------------------------------------------------------------------
def _gate_of(op):
    return getattr(op, "gate", None)


def _is_measurement(op) -> bool:
    g = _gate_of(op)
    return g is not None and type(g).__name__ == "MeasurementGate"


def _is_pauli_string(op) -> bool:
    return type(op).__name__ in ("PauliString", "SingleQubitPauliStringGateOperation") and hasattr(op, "items")


def _has_channel(op) -> bool:
    g = _gate_of(op)
    obj = g if g is not None else op
    if getattr(obj, "_has_unitary_", lambda: False)() or is_parameterized(obj):
        return False
    return hasattr(obj, "_kraus_") or hasattr(obj, "_mixture_")


def kraus_of(obj) -> list[np.ndarray]:
    """Kraus operators: _kraus_, else sqrt(p) U of _mixture_ (cirq-core/cirq/protocols/kraus_protocol.py:138-230)."""
    f = getattr(obj, "_kraus_", None)
    if f is not None:
        r = f()
        if r is not NotImplemented and r is not None:
            return [np.asarray(k, dtype=np.complex128) for k in r]
    f = getattr(obj, "_mixture_", None)
    if f is not None:
        r = f()
        if r is not NotImplemented and r is not None:
            return [np.sqrt(p) * np.asarray(u, dtype=np.complex128) for p, u in r]
    u = getattr(obj, "_unitary_", lambda: NotImplemented)()
    if u is not NotImplemented and u is not None:
        return [np.asarray(u, dtype=np.complex128)]
    raise UnsupportedOperation(f"{obj!r} has no unitary, mixture or Kraus representation")


def _unitary_batch_fn(gate):
    """Returns (exprs, fn(vals)->(B,d,d)) for a gate, or raises UnsupportedOperation."""
    from .gates import Gate

    exprs = tuple(gate._param_exprs()) if hasattr(gate, "_param_exprs") else ()
    mb = getattr(type(gate), "_matrix_batch", None)
    if mb is not None and mb is not Gate._matrix_batch:
        return exprs, gate._matrix_batch
    try:
        u = getattr(gate, "_unitary_", lambda: NotImplemented)()
    except NotImplementedError:
        u = NotImplemented
    if u is NotImplemented or u is None:
        raise UnsupportedOperation(f"{gate!r} has no unitary")
    u = np.asarray(u, dtype=np.complex128)
    return (), (lambda vals, u=u: u[None])


class Lowering:
    """Lowers a circuit into LOps for a state vector (dm=False) or density matrix (dm=True)."""

    def __init__(self, qubits: Sequence, dm: bool, *, split_dm_diag: bool = False):
        self.qubits = tuple(qubits)
        self.index = {q: i for i, q in enumerate(self.qubits)}
        self.n = len(self.qubits)
        self.dm = dm
        # density matrix: a diagonal gate as separate row (d) and column (conj d) diagonals, each
        # linear in the gate's matrix (what gradients need), instead of one d (x) conj(d) diagonal
        self.split_dm_diag = split_dm_diag
        self.ops: list[LOp] = []
        self.nonunitary: list[int] = []  # indices into ops of channels and non-unit scale factors
        self.measurements: list[Measurement] = []
        self.param_refs: list[ParamRef] = []
        self.n_slots = 0

    # bit helpers
    def row_bit(self, q) -> int:
        i = self.index[q]
        return (2 * self.n - 1 - i) if self.dm else (self.n - 1 - i)

    def col_bit(self, q) -> int:
        return self.n - 1 - self.index[q]

    def lower_circuit(self, circuit, *, measurements_terminal_only=True):
        moments = list(circuit)
        for mi, moment in enumerate(moments):
            for op in moment:
                self.lower_op(op)

    def lower_op(self, op):
        if _is_measurement(op):
            g = _gate_of(op)
            self.measurements.append(
                Measurement(g.key, tuple(op.qubits), tuple(getattr(g, "full_invert_mask", lambda: ())()))
            )
            return
        if _is_pauli_string(op) and _gate_of(op) is None:
            self._lower_pauli_string(op)
            return
        from .circuits import _decomposed, _needs_decompose

        if _needs_decompose(op):  # sub-circuits, ops that only decompose, wide decomposable gates
            subs = _decomposed(op)
            if subs != [op]:
                for sub in subs:
                    self.lower_op(sub)
                return
        g = _gate_of(op)
        if g is None:
            raise UnsupportedOperation(f"cannot simulate {op!r}: not a gate operation")
        if _has_channel(op):
            if not self.dm:
                raise UnsupportedOperation(
                    f"{op!r} is a channel; use DensityMatrixSimulator (trajectory sampling of channels on "
                    "state vectors is not supported)"
                )
            self._lower_channel(g, tuple(op.qubits))
            return
        self._lower_gate(g, tuple(op.qubits))

    def _lower_pauli_string(self, ps):
        from .paulis import PauliString

        coef = ps.coefficient
        if isinstance(coef, sympy.Basic):
            raise UnsupportedOperation("PauliString operations with symbolic coefficients are not supported")
        for q, p in ps.items():
            self._lower_gate(p, (q,))
        if complex(coef) != 1:
            self._add_scale(complex(coef))

    def _add_scale(self, c: complex):
        if self.dm:
            c = abs(c) ** 2
        if abs(abs(c) - 1) > 1e-12:
            self.nonunitary.append(len(self.ops))
        self.ops.append(LOp("scale", (), np.complex128(c)))

    def _lower_gate(self, gate, qubits):
        nq = len(qubits)
        if nq == 0:
            exprs, fn = _unitary_batch_fn(gate)
            if any(is_parameterized(e) for e in exprs):
                raise UnsupportedOperation("parameterised global phases are not supported")
            self._add_scale(complex(fn(tuple(np.array([float(e)]) for e in exprs))[0, 0, 0]))
            return
        controls = ()
        sub = gate
        tq = qubits
        if type(gate).__name__ == "ControlledGate":
            nc = gate.num_controls
            cv = tuple(gate.control_values)
            controls = tuple(zip(qubits[:nc], cv))
            sub = gate.sub_gate
            tq = qubits[nc:]
        exprs, fn = _unitary_batch_fn(sub)
        param = any(is_parameterized(e) for e in exprs)
        diagonal = bool(getattr(sub, "_diagonal", False))
        if not param:
            m = fn(tuple(np.array([float(e)]) for e in exprs))[0]
            if not diagonal and np.count_nonzero(m - np.diag(np.diag(m))) == 0:
                diagonal = True
        else:
            m = None
        ref = None
        if param:
            ref = ParamRef(sub, lambda b: b, exprs, slot=self.n_slots)
            self.n_slots += len(exprs)
            self.param_refs.append(ref)
        if diagonal and len(controls) + len(tq) <= (MAX_DIAG // 2 if self.dm else MAX_DIAG):
            # a controlled diagonal is a diagonal on controls + targets
            self._emit_diag(tq, controls, m, ref)
            return
        if len(tq) > MAX_DENSE:
            raise UnsupportedOperation(f"{gate!r}: dense gates on more than {MAX_DENSE} qubits are not supported")
        rows = tuple(self.row_bit(q) for q in reversed(tq))
        rctl = tuple((self.row_bit(q), v) for q, v in controls)
        self.ops.append(LOp("dense", rows, m, rctl, _sub_ref(ref, lambda b: b)))
        if self.dm:
            cols = tuple(self.col_bit(q) for q in reversed(tq))
            cctl = tuple((self.col_bit(q), v) for q, v in controls)
            self.ops.append(
                LOp("dense", cols, None if m is None else np.conj(m), cctl, _sub_ref(ref, np.conj))
            )

    def _emit_diag(self, tq, controls, m, ref):
        """Diagonal of (controls + targets); controls not satisfied give 1."""
        cq = tuple(q for q, _ in controls)
        cv = tuple(v for _, v in controls)
        allq = cq + tuple(tq)
        k = len(allq)
        nt = len(tq)

        def expand(d):  # (B, 2^nt) -> (B, 2^k) big-endian over allq
            B = d.shape[0]
            out = np.ones((B, 2**k), dtype=np.complex128)
            ctl = 0
            for v in cv:
                ctl = ctl * 2 + v
            out[:, ctl * 2**nt : (ctl + 1) * 2**nt] = d
            return out

        def to_diag(b):
            return expand(np.diagonal(b, axis1=-2, axis2=-1))

        d = None if m is None else to_diag(m[None])[0]
        # pure CZPow-like phase on two qubits: diag(1, 1, 1, phi)
        if self.dm and self.split_dm_diag:
            rbits = tuple(self.row_bit(q) for q in reversed(allq))
            cbits = tuple(self.col_bit(q) for q in reversed(allq))
            self.ops.append(LOp("diag", rbits, d, (), _sub_ref(ref, to_diag)))
            self.ops.append(
                LOp("diag", cbits, None if d is None else np.conj(d), (), _sub_ref(ref, lambda b: np.conj(to_diag(b))))
            )
            return
        if self.dm:
            rbits = tuple(self.row_bit(q) for q in reversed(allq))
            cbits = tuple(self.col_bit(q) for q in reversed(allq))

            def dm_diag(dd):  # (B, 2^k) -> (B, 4^k): index = r * 2^k + c over bits (cbits, rbits)
                return (dd[:, :, None] * np.conj(dd[:, None, :])).reshape(dd.shape[0], -1)

            payload = None if d is None else dm_diag(d[None])[0]
            self.ops.append(
                LOp("diag", cbits + rbits, payload, (), _sub_ref(ref, lambda b: dm_diag(to_diag(b))))
            )
            return
        bits = tuple(self.row_bit(q) for q in reversed(allq))
        if d is not None and k == 2 and np.allclose(d[:3], 1, atol=0, rtol=0):
            self.ops.append(LOp("pair", bits, np.complex128(d[3])))
            return
        self.ops.append(LOp("diag", bits, d, (), _sub_ref(ref, to_diag)))

    def _lower_channel(self, g, qubits):
        ks = kraus_of(g)
        k = len(qubits)
        if k > 2:
            raise UnsupportedOperation(f"{g!r}: channels on more than 2 qubits are not supported")
        sop = sum(np.kron(K, np.conj(K)) for K in ks)
        cols = tuple(self.col_bit(q) for q in reversed(qubits))
        rows = tuple(self.row_bit(q) for q in reversed(qubits))
        self.nonunitary.append(len(self.ops))
        self.ops.append(LOp("dense", cols + rows, sop))


def _sub_ref(ref: ParamRef | None, transform, deriv: int | None = None):
    if ref is None:
        return None
    base = ref.transform
    return ParamRef(ref.gate, lambda b, base=base: transform(base(b)), ref.exprs, ref.slot,
                    ref.deriv if deriv is None else deriv)


# --------------------------------------------------------------------------------------------
def _as_matrix(op: LOp) -> np.ndarray:
    if op.kind == "dense":
        return op.mat
    if op.kind == "diag":
        return np.diag(op.mat)
    raise ValueError(op.kind)


def fuse_single_bit(ops: list[LOp]) -> list[LOp]:
    """Multiplies runs of fixed one-bit ops (dense or diagonal) on the same bit into one op.

    Exact matrix products in complex128, as merge_k_qubit_unitaries does
    (cirq-core/cirq/transformers/merge_k_qubit_gates.py:55-63).
    """
    out: list[LOp] = []
    pending: dict[int, int] = {}
    for op in ops:
        if op.fixed and op.kind in ("dense", "diag") and len(op.targets) == 1 and not op.controls:
            b = op.targets[0]
            if b in pending:
                prev = out[pending[b]]
                if prev.kind == "diag" and op.kind == "diag":
                    prev.mat = op.mat * prev.mat
                else:
                    prev.mat = _as_matrix(op) @ _as_matrix(prev)
                    prev.kind = "dense"
                continue
            out.append(LOp(op.kind, op.targets, np.array(op.mat), (), None))
            pending[b] = len(out) - 1
            continue
        for b in op.bits:
            pending.pop(b, None)
        out.append(op)
    return out


# --------------------------------------------------------------------------------------------
# Block fusion inside a pass.  The reference's analogue is merge_k_qubit_unitaries
# (cirq-core/cirq/transformers/merge_k_qubit_gates.py:29-125): connected runs of gates on at most
# k qubits become one matrix, equal to the exact product of the originals.

def embed(M: np.ndarray, tbits, ctrls, B) -> np.ndarray:
    """An op's matrix (little-endian over tbits, controls ((bit, val), ...)) over block bits B."""
    nb = len(B)
    D = 1 << nb
    pos = [B.index(b) for b in tbits]
    k = len(pos)
    cols = np.arange(D, dtype=np.int64)
    ci = np.zeros(D, dtype=np.int64)
    for m, p in enumerate(pos):
        ci |= ((cols >> p) & 1) << m
    tmask = 0
    for p in pos:
        tmask |= 1 << p
    base = cols & ~tmask
    active = np.ones(D, dtype=bool)
    for b, v in ctrls:
        active &= ((cols >> B.index(b)) & 1) == int(v)
    dep = np.zeros(1 << k, dtype=np.int64)
    for ri in range(1 << k):
        for m, p in enumerate(pos):
            dep[ri] |= ((ri >> m) & 1) << p
    out = np.zeros((D, D), dtype=np.complex128)
    ac = np.nonzero(active)[0]
    rows = base[None, ac] | dep[:, None]
    out[rows, ac[None, :]] = M[:, ci[ac]]
    ia = np.nonzero(~active)[0]
    out[ia, ia] = 1
    return out


def _op_matrix(op: LOp) -> np.ndarray:
    if op.kind == "dense":
        return op.mat
    if op.kind == "diag":
        return np.diag(op.mat)
    if op.kind == "pair":
        return np.diag([1, 1, 1, op.mat]).astype(np.complex128)
    raise ValueError(op.kind)


# Relative costs (arbitrary units) the planners use to choose how a fused group runs.  Marginal
# cost per op measured on MI355X at 28 qubits, tile of 12 bits (ms): register dense1 0.100,
# dense2 0.139, dense3 0.245; through LDS dense1 0.143, dense2 0.176, dense3 0.383; MFMA block
# k=4 0.177, k=5 0.275; PERMUTE 0.058; diagonal 0.107; CZ pair phase 0.033.  The table is not
# proportional to these: its MFMA weights were fit by timing whole-circuit plans (with blocks left
# at the end of a pass merged, the 30-qubit grid circuit and the 16-qubit noisy density matrix
# were fastest around these values), and the autotuner scales them per circuit.
DEFAULT_KMAX = {False: 4, True: 5}  # widest fused block: state vectors, density matrices (timed)
COSTS = {"dense1": 1.0, "dense2": 1.8, "dense3": 1.75, "diag": 0.05, "pair": 0.03, "scale": 0.03, "mfma4": 2.0,
         "mfma5": 2.55, "reg1": 0.45, "reg2": 0.9, "reg3": 1.8, "permute": 0.8}
# Entries replaced for programs that run as specialised kernels (cirq_amd.specialize).  Measured
# there at 30 qubits, 13-bit tiles: a 3-qubit dense op through LDS adds 0.87 ms to its pass (its
# VALU floor, 32 FMAs per amplitude), a 4-qubit MFMA block 0.37 ms; with dense3 above mfma4 the
# planner widens such groups to MFMA blocks, and the 30-qubit grid program's passes take 0.8% less.
SPECIALIZED_COSTS = {"dense3": 2.6}


def _valu_cost(op: LOp, costs: dict | None = None) -> float:
    costs = COSTS if costs is None else costs
    if op.kind == "dense":
        return costs[f"dense{len(op.targets)}"] if len(op.targets) <= 3 else 1e9
    return costs.get(op.kind, 0.3)


MAX_VARIANT_BITS = 3  # outer bits a fused op may depend on (2^3 matrices per op)


class _Block:
    __slots__ = ("bits", "obits", "mats", "members")

    def __init__(self, bits, obits, mats, members):
        self.bits = list(bits)  # local bits, matrix index bit m <-> bits[m]
        self.obits = list(obits)  # outer bits; variant index bit i <-> obits[i]
        self.mats = mats  # (2^len(obits), 2^len(bits), 2^len(bits))
        self.members = list(members)


def _variant_matrix(op: LOp, local_bits: list, outer: list, assignment: int) -> np.ndarray:
    """op's action on local_bits when the outer bits `outer` take the values of `assignment`."""
    val = {b: (assignment >> i) & 1 for i, b in enumerate(outer)}
    if op.kind == "dense":
        lc = [(b, v) for b, v in op.controls if b not in val]
        if any(val[b] != int(v) for b, v in op.controls if b in val):
            return np.eye(1 << len(local_bits), dtype=np.complex128)
        return embed(op.mat, list(op.targets), tuple(lc), local_bits)
    if op.kind in ("diag", "pair"):
        if op.kind == "diag":
            tb, dvals = list(op.targets), np.asarray(op.mat)
        else:
            tb, dvals = list(op.targets), np.array([1, 1, 1, op.mat], dtype=np.complex128)
        lt = [b for b in tb if b not in val]
        sub = np.empty(1 << len(lt), dtype=np.complex128)
        for x in range(1 << len(lt)):
            idx, li = 0, 0
            for m, b in enumerate(tb):
                if b in val:
                    bit = val[b]
                else:
                    bit = (x >> li) & 1
                    li += 1
                idx |= bit << m
            sub[x] = dvals[idx]
        return embed(np.diag(sub), lt, (), local_bits)
    raise ValueError(op.kind)


def _rebase(blk: _Block, bits: list, obits: list) -> np.ndarray:
    """blk's matrices over (bits, obits): embedded on the local bits, repeated over new outer bits."""
    out = np.empty((1 << len(obits), 1 << len(bits), 1 << len(bits)), dtype=np.complex128)
    for a in range(1 << len(obits)):
        sub = 0
        for i, b in enumerate(blk.obits):
            sub |= ((a >> obits.index(b)) & 1) << i
        out[a] = embed(blk.mats[sub], blk.bits, (), bits)
    return out


def fuse_blocks(ops: list[LOp], local: set, L: int, kmax: int = 5, mfma: bool = True, *, costs: dict | None = None,
                variant_bits: int | None = None, kron_merge: bool = True) -> list[LOp]:
    """Fuses a pass's fixed ops on local bits into blocks of at most kmax local bits.

    Several blocks stay open at once (one per disjoint bit group), so a block keeps absorbing
    gates on its bits across many cycles until an op links it to bits it cannot take.  Ops that
    also touch bits outside the tile (a CZ to an outer qubit, a gate controlled by one) are
    absorbed too: the block then holds one matrix per value of up to MAX_VARIANT_BITS outer bits,
    and each tile applies the one its outer bits select.  A closed block runs as: its original
    ops, one fused dense op (k <= 3, VALU) or an MFMA block (k = 4, 5; smaller groups padded),
    whichever the cost table rates cheapest (`costs`, default COSTS; `variant_bits` caps the outer
    bits, default MAX_VARIANT_BITS).  With kron_merge, small blocks still open at the end of the
    pass share MFMA blocks (Kronecker products of disjoint blocks) when that rates cheaper."""
    costs = COSTS if costs is None else costs
    vmax = MAX_VARIANT_BITS if variant_bits is None else min(int(variant_bits), MAX_VARIANT_BITS)
    out: list[LOp] = []
    open_blocks: list[_Block] = []
    owner: dict = {}
    use_mfma = mfma and L >= 9

    def close(blk: _Block):
        open_blocks.remove(blk)
        for b in blk.bits:
            owner.pop(b, None)
        members = blk.members
        if len(members) == 1:
            out.append(members[0])
            return
        nb = len(blk.bits)
        separate = sum(_valu_cost(m, costs) for m in members)
        fused = costs[f"dense{nb}"] if nb <= 3 else 1e9
        mf = (costs["mfma4"] if nb <= 4 else costs["mfma5"]) if use_mfma else 1e9
        best = min(separate, fused, mf)
        vb = tuple(blk.obits)
        single = blk.mats[0] if not vb else blk.mats
        if best == separate:
            out.extend(members)
        elif best == fused:
            out.append(LOp("dense", tuple(blk.bits), single, vbits=vb))
        else:
            bits = list(blk.bits)
            mats = blk.mats
            if nb < 4:
                extra = [b for b in sorted(local) if b not in bits][: 4 - nb]
                mats = np.stack([embed(m, bits, (), bits + extra) for m in mats])
                bits = bits + extra
            out.append(LOp("block", tuple(bits), mats[0] if not vb else mats, vbits=vb))

    for op in ops:
        fixed = op.fixed and op.kind in ("dense", "diag", "pair") and not op.vbits
        bits = list(dict.fromkeys(list(op.targets) + [b for b, _ in op.controls]))
        lb = [b for b in bits if b in local]
        ob = [b for b in bits if b not in local]
        absorbable = fixed and lb and len(lb) <= kmax and len(ob) <= vmax
        if absorbable and op.kind == "dense" and any(b not in local for b in op.targets):
            absorbable = False
        if not absorbable:
            for blk in {id(owner[b]): owner[b] for b in op.bits if b in owner}.values():
                close(blk)
            out.append(op)
            continue
        touched = []
        for b in lb:
            blk = owner.get(b)
            if blk is not None and blk not in touched:
                touched.append(blk)
        ubits = []
        uobits = []
        for blk in touched:
            ubits += blk.bits
            uobits += [b for b in blk.obits if b not in uobits]
        ubits += [b for b in lb if b not in ubits]
        uobits += [b for b in ob if b not in uobits]
        if len(ubits) > kmax or len(uobits) > vmax:
            for blk in touched:
                close(blk)
            touched = []
            ubits, uobits = list(lb), list(ob)
        mats = np.broadcast_to(np.eye(1 << len(ubits), dtype=np.complex128),
                               (1 << len(uobits), 1 << len(ubits), 1 << len(ubits))).copy()
        members = []
        for blk in touched:
            mats = np.einsum("vij,vjk->vik", _rebase(blk, ubits, uobits), mats)
            members += blk.members
            open_blocks.remove(blk)
        for a in range(1 << len(uobits)):
            mats[a] = _variant_matrix(op, ubits, uobits, a) @ mats[a]
        nb = _Block(ubits, uobits, mats, members + [op])
        open_blocks.append(nb)
        for b in ubits:
            owner[b] = nb
    # End of the pass: every open block closes and none can grow any more.  Blocks that would run
    # as VALU ops (a few one- or two-qubit gates each) are disjoint and nothing follows them in the
    # pass, so several can share one MFMA block: the Kronecker product of their matrices.
    small = []
    for blk in list(open_blocks):
        nb = len(blk.bits)
        own = min(sum(_valu_cost(m, costs) for m in blk.members), costs[f"dense{nb}"] if nb <= 3 else 1e9)
        mf = (costs["mfma4"] if nb <= 4 else costs["mfma5"]) if use_mfma else 1e9
        if use_mfma and kron_merge and nb <= 3 and own < mf:
            small.append((own, blk))
        else:
            close(blk)
    small.sort(key=lambda x: (-len(x[1].bits), -x[0]))
    groups: list[list] = []
    for own, blk in small:
        for g in groups:
            bits = sum(len(b.bits) for _, b in g) + len(blk.bits)
            obits = set(ob for _, b in g for ob in b.obits) | set(blk.obits)
            if bits <= kmax and len(obits) <= vmax:
                g.append((own, blk))
                break
        else:
            groups.append([(own, blk)])
    for g in groups:
        nb = sum(len(b.bits) for _, b in g)
        mf = costs["mfma4"] if nb <= 4 else costs["mfma5"]
        if len(g) < 2 or mf >= sum(own for own, _ in g):
            for _, blk in g:
                close(blk)
            continue
        bits = [x for _, b in g for x in b.bits]
        obits = list(dict.fromkeys(ob for _, b in g for ob in b.obits))
        mats = np.broadcast_to(np.eye(1 << nb, dtype=np.complex128), (1 << len(obits), 1 << nb, 1 << nb)).copy()
        for _, blk in g:
            mats = np.einsum("vij,vjk->vik", _rebase(blk, bits, obits), mats)
            open_blocks.remove(blk)
            for b in blk.bits:
                owner.pop(b, None)
        if nb < 4:
            extra = [b for b in sorted(local) if b not in bits][: 4 - nb]
            mats = np.stack([embed(m, bits, (), bits + extra) for m in mats])
            bits = bits + extra
        vb = tuple(obits)
        out.append(LOp("block", tuple(bits), mats[0] if not vb else mats, vbits=vb))
    return out


def block_payload(U: np.ndarray) -> np.ndarray:
    """MFMA A operands of a 2^k block (k = 4, 5): real form split into FP16 hi + lo, in
    [part][ks][rb][lane][8] order (see CQS_OP_BLOCK), returned as complex64 slots."""
    D = U.shape[0]
    R = 2 * D
    Rm = np.zeros((R, R))
    Rm[0::2, 0::2] = U.real
    Rm[0::2, 1::2] = -U.imag
    Rm[1::2, 0::2] = U.imag
    Rm[1::2, 1::2] = U.real
    hi = Rm.astype(np.float16)
    lo = (Rm - hi.astype(np.float64)).astype(np.float16)
    KS, RB = R // 16, R // 32
    lane = np.arange(64)
    out = np.zeros((2, KS, RB, 64, 8), dtype=np.float16)
    for part, M in enumerate((hi, lo)):
        for ks in range(KS):
            for rb in range(RB):
                rows = rb * 32 + lane % 32
                cols = ks * 16 + 8 * (lane // 32)[:, None] + np.arange(8)[None, :]
                out[part, ks, rb] = M[rows[:, None], cols]
    return np.frombuffer(out.tobytes(), dtype=np.complex64)


def block_matrix_from_payload(payload: np.ndarray, k: int) -> np.ndarray:
    """Inverse of block_payload (hi + lo in float64), for tests."""
    D = 1 << k
    R = 2 * D
    KS, RB = R // 16, R // 32
    arr = np.frombuffer(np.ascontiguousarray(payload).tobytes(), dtype=np.float16).reshape(2, KS, RB, 64, 8)
    Rm = np.zeros((R, R))
    lane = np.arange(64)
    for part in range(2):
        for ks in range(KS):
            for rb in range(RB):
                rows = rb * 32 + lane % 32
                cols = ks * 16 + 8 * (lane // 32)[:, None] + np.arange(8)[None, :]
                Rm[rows[:, None], cols] += arr[part, ks, rb].astype(np.float64)
    return Rm[0::2, 0::2] + 1j * Rm[1::2, 0::2]


# --------------------------------------------------------------------------------------------
@dataclasses.dataclass
class Pass:
    L: int
    lpos: tuple  # ascending flat bit positions
    ops: list


def schedule(ops: list[LOp], n_bits: int, L: int, coalesce: int) -> list[Pass]:
    """Greedy pass construction.

    Each pass starts with the `coalesce` lowest bits local (so every tile row is a contiguous run
    of 2^coalesce amplitudes) and walks the remaining ops in order: a dense op joins the pass if
    its targets fit into the L local bits, diagonal ops always join; an op that cannot join blocks
    its bits for the rest of the pass so that per-bit order is preserved (ops on disjoint bits
    commute).  A fence op (one that does not commute with earlier ops on other bits, such as a
    channel adjoint applied to one of the two states of an adjoint sweep) joins only when no
    earlier op was deferred, and once deferred it holds back every later op.
    """
    L = min(L, n_bits)
    c = min(coalesce, L)
    passes: list[Pass] = []
    rest = list(ops)

    def one_pass(forced: set):
        S = set(forced)
        blocked: set = set()
        take, defer = [], []
        fenced = False
        for op in rest:
            bits = op.bits
            if fenced or (op.fence and defer):
                defer.append(op)
                fenced = True
                continue
            if bits & blocked:
                defer.append(op)
                blocked |= bits
                continue
            if op.kind in ("dense", "grad_dense"):
                need = (set(op.targets) | set(op.need)) - S
                if len(S) + len(need) <= L:
                    S |= need
                    take.append(op)
                else:
                    defer.append(op)
                    blocked |= bits
                    fenced = op.fence
            else:
                take.append(op)
        return S, take, defer

    while rest:
        S, take, defer = one_pass(set(range(c)))
        if not take:  # the first op is too wide to share the tile with the coalescing bits
            S, take, defer = one_pass(set())
        if not take:
            raise UnsupportedOperation("scheduler made no progress (op wider than the tile?)")
        # fill the tile with the lowest unused bits
        b = 0
        while len(S) < L:
            if b not in S:
                S.add(b)
            b += 1
        passes.append(Pass(L, tuple(sorted(S)), take))
        rest = defer
    return passes


# --------------------------------------------------------------------------------------------
def variant_code(vbits) -> int:
    """t[5] of a DENSE / BLOCK op whose matrix depends on outer bits: count << 24 | positions, 7
    bits each; the op's matrices follow each other at `mat`, variant index bit i = flat bit
    vbits[i] of the tile's outer index."""
    vbits = tuple(int(b) for b in vbits)
    if len(vbits) > MAX_VARIANT_BITS or any(not 0 <= b < 128 for b in vbits):
        raise ValueError(f"bad variant bits {vbits}")
    code = len(vbits) << 24
    for i, b in enumerate(vbits):
        code |= b << (7 * i)
    return code


class PackedProgram:
    """Host/device arrays of a scheduled program in the C ABI layout."""

    def __init__(self, passes: list[Pass], n_bits: int):
        self.n_bits = n_bits
        recs = []
        mats: list[np.ndarray] = []
        aux: list[int] = []
        self.param_entries: list[tuple[int, ParamRef, int]] = []  # (mats offset, ref, size)
        self.pass_arr = np.zeros(len(passes), dtype=_lib.PASS_DTYPE)
        mat_off = 0

        def add_mat(arr) -> int:
            nonlocal mat_off
            a = np.asarray(arr, dtype=np.complex64).reshape(-1)
            mats.append(a)
            o = mat_off
            mat_off += a.size
            return o

        for pi, ps in enumerate(self.passes_merged(passes)):
            local = {b: j for j, b in enumerate(ps.lpos)}
            layout = list(ps.lpos)
            begin = len(recs)
            max_k = 0
            for op in ps.ops:
                r = np.zeros((), dtype=_lib.OP_DTYPE)
                r["slot"] = -1
                if op.kind == "permute":
                    new_local = {b: j for j, b in enumerate(op.targets)}
                    if set(new_local) != set(local):
                        raise ValueError("permute must keep the pass's local bits")
                    r["kind"] = _lib.OP_PERMUTE
                    r["k"] = ps.L
                    r["t"][0] = len(aux)
                    dst = [new_local[layout[b]] for b in range(ps.L)]
                    if sorted(dst) != list(range(ps.L)):  # the kernel writes LDS at these indices
                        raise ValueError("PERMUTE destinations must be a permutation of the tile bits")
                    aux.extend(dst)
                    local, layout = new_local, list(op.targets)
                    recs.append(r)
                    continue
                if (op.kind == "dense" and len(op.targets) <= 3) or op.kind == "block":
                    op = sorted_targets(op, key=local.get)
                if op.kind == "dense":
                    local_ctl = [(b, v) for b, v in op.controls if b in local]
                    outer_ctl = [(b, v) for b, v in op.controls if b not in local]
                    com = cov = 0
                    for b, v in outer_ctl:
                        com |= 1 << b
                        cov |= int(v) << b
                    r["co_mask_lo"], r["co_mask_hi"] = com & 0xFFFFFFFF, com >> 32
                    r["co_val_lo"], r["co_val_hi"] = cov & 0xFFFFFFFF, cov >> 32
                    k = len(op.targets)
                    if k + (len(local_ctl) if k >= 4 else 0) > 5:
                        raise UnsupportedOperation("dense op on more than 5 tile bits")
                    if k <= 3:
                        r["kind"] = _lib.OP_DENSE
                        r["k"] = k
                        max_k = max(max_k, k)
                        for m, b in enumerate(op.targets):
                            r["t"][m] = local[b]
                        clm = clv = 0
                        for b, v in local_ctl:
                            clm |= 1 << local[b]
                            clv |= int(v) << local[b]
                        r["cl_mask"], r["cl_val"] = clm, clv
                        d = 2**k
                        payload = op.mat if op.mat is not None else np.zeros((d, d))
                        want = ((1 << len(op.vbits)), d, d) if op.vbits else (d, d)
                        if np.shape(payload) != want:  # the kernel reads exactly this many matrices
                            raise ValueError(f"dense payload {np.shape(payload)} != {want}")
                        r["mat"] = add_mat(payload)  # variants, if any, one after the other
                        r["t"][5] = variant_code(op.vbits)
                        if op.pref is not None:
                            self.param_entries.append((int(r["mat"]), op.pref, d * d))
                            r["slot"] = op.pref.slot
                    else:
                        # 4-5 bit matrices run on MFMA; local controls become part of the matrix
                        bits = list(op.targets) + [b for b, _ in local_ctl]
                        r["kind"] = _lib.OP_BLOCK
                        r["k"] = len(bits)
                        for m, b in enumerate(bits):
                            r["t"][m] = local[b]
                        tb, lc = list(op.targets), tuple(local_ctl)
                        to_block = lambda M, tb=tb, lc=lc, bits=bits: embed(M, tb, lc, bits)
                        if mat_off % 2:
                            add_mat(np.zeros(1))
                        d = 2 ** len(bits)
                        payload = op.mat if op.mat is not None else np.eye(d)
                        r["mat"] = add_mat(block_payload(to_block(payload) if op.mat is not None else payload))
                        if op.pref is not None:
                            base_tf = op.pref.transform
                            ref = ParamRef(
                                op.pref.gate,
                                lambda b, base_tf=base_tf, to_block=to_block: np.stack(
                                    [block_payload(to_block(x)) for x in base_tf(b)]
                                ),
                                op.pref.exprs,
                                op.pref.slot,
                            )
                            self.param_entries.append((int(r["mat"]), ref, 2 * d * d))
                            r["slot"] = op.pref.slot
                elif op.kind == "block":
                    k = len(op.targets)
                    r["kind"] = _lib.OP_BLOCK
                    r["k"] = k
                    for m, b in enumerate(op.targets):
                        r["t"][m] = local[b]
                    if mat_off % 2:  # 16-byte alignment of the half8 operands
                        add_mat(np.zeros(1))
                    variants = op.mat if op.vbits else [op.mat]
                    if len(variants) != 1 << len(op.vbits) or np.shape(variants[0]) != (1 << k, 1 << k):
                        raise ValueError("block payload does not match its variant bits")
                    r["mat"] = add_mat(np.concatenate([block_payload(m) for m in variants]))
                    r["t"][5] = variant_code(op.vbits)
                elif op.kind == "diag":
                    k = len(op.targets)
                    r["kind"] = _lib.OP_DIAG
                    r["k"] = k
                    for m, b in enumerate(op.targets):
                        r["t"][m] = local[b] if b in local else OUTER + b
                    payload = op.mat if op.mat is not None else np.zeros(2**k)
                    r["mat"] = add_mat(payload)
                    if op.pref is not None:
                        self.param_entries.append((int(r["mat"]), op.pref, 2**k))
                        r["slot"] = op.pref.slot
                elif op.kind == "pairs":
                    r["kind"] = _lib.OP_PAIRPHASE
                    r["k"] = len(op.targets) // 2
                    r["t"][0] = len(aux)
                    aux.extend(local[b] if b in local else OUTER + int(b) for b in op.targets)
                    r["mat"] = add_mat(op.mat)
                elif op.kind == "scale":
                    r["kind"] = _lib.OP_SCALE
                    r["mat"] = add_mat(np.array([op.mat]))
                elif op.kind == "grad_dense":
                    k = len(op.targets)
                    if k > 3:
                        raise UnsupportedOperation("gradients of gates on more than 3 qubits are not supported")
                    r["kind"] = _lib.OP_GRAD_DENSE
                    r["k"] = k
                    for m, b in enumerate(op.targets):
                        r["t"][m] = local[b]
                    d = 2**k
                    r["mat"] = add_mat(np.zeros((d, d)))
                    self.param_entries.append((int(r["mat"]), op.pref, d * d))
                    r["slot"] = op.slot
                elif op.kind == "grad_diag":
                    k = len(op.targets)
                    r["kind"] = _lib.OP_GRAD_DIAG
                    r["k"] = k
                    for m, b in enumerate(op.targets):
                        r["t"][m] = local[b] if b in local else OUTER + b
                    r["mat"] = add_mat(np.zeros(2**k))
                    self.param_entries.append((int(r["mat"]), op.pref, 2**k))
                    r["slot"] = op.slot
                else:
                    raise ValueError(op.kind)
                if op.sel:
                    r["kind"] = int(r["kind"]) | (op.sel << 8)
                recs.append(r)
            p = self.pass_arr[pi]
            p["L"] = ps.L
            p["op_begin"] = begin
            p["op_count"] = len(recs) - begin
            p["max_k"] = max_k
            p["lpos"][: ps.L] = ps.lpos
            p["lpos_out"][: ps.L] = layout
        self.ops_arr = np.array(recs, dtype=_lib.OP_DTYPE) if recs else np.zeros(1, dtype=_lib.OP_DTYPE)
        self.mats_arr = np.concatenate(mats) if mats else np.zeros(1, dtype=np.complex64)
        self.aux_arr = np.array(aux if aux else [0], dtype=np.int32)
        self.n_passes = len(self.pass_arr)
        self.n_ops = len(recs)

    @staticmethod
    def passes_merged(passes: list[Pass]) -> list[Pass]:
        """Merges every run of elementwise ops' 'pair' phases into one PAIRPHASE op."""
        out = []
        for ps in passes:
            ops: list[LOp] = []
            run_pairs: list[LOp] = []

            def flush():
                for i in range(0, len(run_pairs), 8):  # the kernel walks pairs in unrolled chunks of 8
                    chunk = run_pairs[i : i + 8]
                    bits = tuple(b for p in chunk for b in p.targets)
                    phases = np.array([p.mat for p in chunk], dtype=np.complex128)
                    ops.append(LOp("pairs", bits, phases))
                run_pairs.clear()

            for op in ps.ops:
                if op.kind == "pair":
                    run_pairs.append(op)
                elif op.kind in ("dense", "block", "grad_dense", "grad_diag") or op.sel != 0:
                    flush()
                    ops.append(op)
                else:
                    ops.append(op)  # elementwise ops commute with the pending pair phases
            flush()
            out.append(Pass(ps.L, ps.lpos, ops))
        return out


def threads_for_tile(L: int) -> int:
    """Workgroup size of the pass kernel for a tile of 2^L amplitudes."""
    return 256 if L <= 12 else (1 << L) // 16


def order_tile_bits(p: Pass, coalesce: int) -> Pass:
    """Orders a pass's local bits: the coalescing bits first (thread lanes, contiguous in HBM),
    then the rest by how many 1-2 bit dense ops use them, busiest last: the top
    log2(2^L / threads) tile bits live in one thread's registers, where those ops skip LDS."""
    L = p.L
    c = min(coalesce, L)
    low = [b for b in range(c) if b in p.lpos]
    rest = [b for b in p.lpos if b not in low]
    use = {b: 0 for b in rest}
    for op in p.ops:
        if op.kind == "dense" and len(op.targets) <= 2 and not any(b in use for b, _ in op.controls):
            for b in op.targets:
                if b in use:
                    use[b] += 1
    rest.sort(key=lambda b: (use[b], -b))
    return Pass(L, tuple(sorted(low) + rest), p.ops)


def sorted_targets(op: LOp, key=None) -> LOp:
    """The same dense op or block with its targets in ascending order of key(bit) (default: the
    bit; the packer passes the bit's tile position), matrix indices permuted to match: the form
    the pass kernel runs in registers."""
    t = list(op.targets)
    k = len(t)
    key = key or (lambda b: b)
    order = sorted(range(k), key=lambda m: key(t[m]))  # new target m' is old target order[m']
    if order == list(range(k)):
        return op
    P = np.array([sum(((i >> m) & 1) << order[m] for m in range(k)) for i in range(1 << k)])
    mat = None if op.mat is None else np.asarray(op.mat)[..., P, :][..., :, P]
    ref = op.pref
    if ref is not None:
        base = ref.transform
        ref = ParamRef(ref.gate, lambda b, base=base, P=P: base(b)[..., P, :][..., :, P], ref.exprs, ref.slot,
                       ref.deriv)
    return LOp(op.kind, tuple(t[m] for m in order), mat, op.controls, ref, op.sel, op.slot, op.vbits)


_ELEMENTWISE = ("diag", "pair", "pairs", "scale")


def _op_dependencies(ops: list[LOp]) -> list[list[int]]:
    """succ[i]: later ops that must follow op i (they share a bit and not both are elementwise)."""
    succ: list[list[int]] = [[] for _ in ops]
    last_mix: dict = {}
    ew_since: dict = {}
    for j, op in enumerate(ops):
        ew = op.kind in _ELEMENTWISE
        deps = set()
        for b in op.bits:
            if b in last_mix:
                deps.add(last_mix[b])
            if not ew:
                deps.update(ew_since.get(b, ()))
        for i in deps:
            succ[i].append(j)
        for b in op.bits:
            if ew:
                ew_since.setdefault(b, []).append(j)
            else:
                last_mix[b] = j
                ew_since[b] = []
    return succ


def plan_registers(p: Pass, coalesce: int, costs: dict | None = None) -> Pass:
    """Orders a pass's ops and inserts PERMUTE ops so that dense ops on 1-3 bits run in registers.

    The pass kernel keeps each thread's 2^NR tile elements in registers (NR = the top tile bits);
    dense ops whose targets are all register bits run there, other dense ops and MFMA blocks
    exchange the tile through LDS.  Ops are list-scheduled over their dependency graph (ops on
    disjoint bits commute, elementwise ops commute with each other): ready elementwise and
    register ops run first; when only LDS ops are ready, the planner either re-lays the tile
    (PERMUTE) so that a new set of register bits lets more ops run in registers, or runs the
    ready op through LDS, whichever the cost table says is cheaper.  The coalescing bits stay the
    lowest tile bits throughout, so loads and stores remain contiguous."""
    costs = COSTS if costs is None else costs
    L = p.L
    NT = threads_for_tile(L)
    TB = NT.bit_length() - 1
    NR = L - TB
    if NR <= 0 or not p.ops:
        return p
    c = min(coalesce, L)
    low = [b for b in p.lpos[:c]]
    movable = [b for b in p.lpos if b not in low]
    ops = p.ops
    n_ops = len(ops)
    succ = _op_dependencies(ops)
    npred = [0] * n_ops
    for i in range(n_ops):
        for j in succ[i]:
            npred[j] += 1
    kmax_reg = min(3, NR)

    def reg_ok(op, regs) -> bool:
        return op.kind == "dense" and len(op.targets) <= kmax_reg and set(op.targets) <= regs

    def lds_cost(op) -> float:
        if op.kind == "dense":
            return costs[f"dense{len(op.targets)}"] if len(op.targets) <= 3 else costs["mfma5"]
        return costs["mfma5"]

    def drain(regs, npred, ready, emit=None) -> float:
        """Runs every elementwise / register op reachable; returns the LDS cost saved."""
        gain = 0.0
        progress = True
        while progress:
            progress = False
            for j in sorted(ready):
                op = ops[j]
                if op.kind in _ELEMENTWISE or reg_ok(op, regs):
                    ready.discard(j)
                    if op.kind == "dense":
                        gain += lds_cost(op) - costs[f"reg{len(op.targets)}"]
                    if emit is not None:
                        emit.append(op)
                    for s2 in succ[j]:
                        npred[s2] -= 1
                        if npred[s2] == 0:
                            ready.add(s2)
                    progress = True
        return gain

    def future_use(npred_now) -> dict:
        use = {b: 0 for b in movable}
        for j in range(n_ops):
            op = ops[j]
            if npred_now[j] >= 0 and op.kind == "dense" and len(op.targets) <= kmax_reg:
                for b in op.targets:
                    if b in use:
                        use[b] += 1
        return use

    def best_regs(npred, ready, done) -> tuple[set, float]:
        alive = [-1 if done[j] else npred[j] for j in range(n_ops)]
        use = future_use(alive)
        S: set = set()
        score = 0.0
        for _ in range(NR):
            best = None
            for b in movable:
                if b in S:
                    continue
                g = drain(S | {b}, list(npred), set(ready))
                key = (g, use[b], -b)
                if best is None or key > best[0]:
                    best = (key, b)
            S.add(best[1])
            score = best[0][0]
        return S, score

    def layout_for(cur: list, S: set) -> list:
        regs_now = cur[TB:]
        keep = [b if b in S else None for b in regs_now]
        newcomers = [b for b in S if b not in regs_now]
        leaving = [b for b in regs_now if b not in S]
        new_regs = []
        for b in keep:
            new_regs.append(b if b is not None else newcomers.pop())
        thread = [b if b not in S else None for b in cur[c:TB]]
        thread = [b if b is not None else leaving.pop() for b in thread]
        return cur[:c] + thread + new_regs

    done = [False] * n_ops
    ready = {j for j in range(n_ops) if npred[j] == 0}
    S, _ = best_regs(npred, ready, done)
    start = low + [b for b in movable if b not in S] + sorted(S)
    layout = list(start)
    out: list[LOp] = []

    index_of = {id(op): j for j, op in enumerate(ops)}

    def mark(op):
        done[index_of[id(op)]] = True

    while True:
        before = len(out)
        drain(set(layout[TB:]), npred, ready, out)
        for op in out[before:]:
            if op.kind != "permute":
                mark(op)
        if not ready:
            break
        # only LDS ops are ready: those that can never run in registers go first
        forced = [j for j in sorted(ready) if not (ops[j].kind == "dense" and len(ops[j].targets) <= kmax_reg
                                                    and not set(ops[j].targets) & set(low))]
        if forced:
            j = forced[0]
        else:
            S, gain = best_regs(npred, ready, done)
            if gain > costs["permute"]:
                layout = layout_for(layout, S)
                out.append(LOp("permute", tuple(layout), None))
                continue
            j = min(ready)
        ready.discard(j)
        out.append(ops[j])
        done[j] = True
        for s2 in succ[j]:
            npred[s2] -= 1
            if npred[s2] == 0:
                ready.add(s2)
    if len([o for o in out if o.kind != "permute"]) != n_ops:
        raise RuntimeError("register planner lost ops")
    return Pass(L, tuple(start), out)


def plan_passes(ops: list[LOp], n_bits: int, *, L: int = 12, coalesce: int = 4, fuse: bool = True,
                blocks: bool = True, kmax: int = 5, costs: dict | None = None,
                variant_bits: int | None = None, kron_merge: bool = True) -> list:
    """The tile passes of a forward program (before packing)."""
    if fuse:
        ops = fuse_single_bit(ops)
    passes = schedule(ops, n_bits, L, coalesce)
    if blocks:
        passes = [Pass(p.L, p.lpos, fuse_blocks(p.ops, set(p.lpos), p.L, kmax, costs=costs, variant_bits=variant_bits,
                                                kron_merge=kron_merge))
                  for p in passes]
    return [plan_registers(order_tile_bits(p, coalesce), coalesce, costs) for p in passes]


def compile_ops(ops: list[LOp], n_bits: int, *, L: int = 12, coalesce: int = 4, fuse: bool = True,
                blocks: bool = True, kmax: int = 5, costs: dict | None = None,
                variant_bits: int | None = None, kron_merge: bool = True) -> PackedProgram:
    return PackedProgram(plan_passes(ops, n_bits, L=L, coalesce=coalesce, fuse=fuse, blocks=blocks, kmax=kmax,
                                     costs=costs, variant_bits=variant_bits, kron_merge=kron_merge), n_bits)


# --------------------------------------------------------------------------------------------
class ParamEvaluator:
    """Evaluates parameter expressions for batches of resolvers (vectorised over resolvers).

    A bare symbol and an affine function of one symbol (the common cases: rx(theta) has exponent
    theta/pi) are evaluated directly; other expressions go through sympy.lambdify over their own
    free symbols only."""

    def __init__(self, refs: Sequence[ParamRef]):
        self.symbols: list[sympy.Symbol] = []
        seen = set()
        for r in refs:
            for e in r.exprs:
                if isinstance(e, sympy.Basic):
                    for s in sorted(e.free_symbols, key=lambda s: s.name):
                        if s.name not in seen:
                            seen.add(s.name)
                            self.symbols.append(s)
        self.index = {s: i for i, s in enumerate(self.symbols)}
        self._fns: dict = {}
        self._dfns: dict = {}

    def _compile(self, e):
        """(kind, data): ('const', v) | ('affine', (i, a, b)) | ('fn', (rows, f))."""
        if not isinstance(e, sympy.Basic) or not e.free_symbols:
            return "const", float(e)
        syms = sorted(e.free_symbols, key=lambda x: x.name)
        rows = [self.index[x] for x in syms]
        if len(syms) == 1:
            a = sympy.diff(e, syms[0])
            if not a.free_symbols:
                return "affine", (rows[0], float(a), float(e.subs(syms[0], 0)))
        return "fn", (rows, sympy.lambdify(syms, e, "numpy"))

    @staticmethod
    def _eval(c, symvals: np.ndarray) -> np.ndarray:
        kind, d = c
        B = symvals.shape[1]
        if kind == "const":
            return np.full(B, d)
        if kind == "affine":
            i, a, b = d
            return a * symvals[i] + b
        rows, f = d
        return np.broadcast_to(np.asarray(f(*(symvals[r] for r in rows)), dtype=np.float64), (B,)).copy()

    def symbol_values(self, resolvers) -> np.ndarray:
        """(n_symbols, B) float64 array of resolved symbol values."""
        B = len(resolvers)
        out = np.zeros((len(self.symbols), B))
        for j, r in enumerate(resolvers):
            for i, s in enumerate(self.symbols):
                v = r.value_of(s)
                if isinstance(v, sympy.Basic):
                    raise ValueError(f"Circuit contains unresolved symbol {s.name!r}")
                out[i, j] = float(np.real(complex(v)))
        return out

    def values(self, e, symvals: np.ndarray) -> np.ndarray:
        c = self._fns.get(e)
        if c is None:
            c = self._fns[e] = self._compile(e)
        return self._eval(c, symvals)

    def dvalues(self, e, sym, symvals: np.ndarray):
        key = (e, sym)
        c = self._dfns.get(key, False)
        if c is False:
            if isinstance(e, sympy.Basic) and sym in e.free_symbols:
                c = self._compile(sympy.diff(e, sym))
            else:
                c = None
            self._dfns[key] = c
        if c is None:
            return None
        return self._eval(c, symvals)


class SlotChain:
    """Chain rule from gradient slots (one per differentiated parameter occurrence) to symbols:
    grad[b, i] = sum over slots s of g[b, s] * d expr_s / d symbol_i at point b.  Derivatives
    that are constants (expressions affine in the symbol, the usual theta / pi) form a sparse
    matrix built once; the others are evaluated per call."""

    def __init__(self, slot_exprs: Sequence, symbols: Sequence, ev: ParamEvaluator):
        index = {sym: i for i, sym in enumerate(symbols)}
        self.n_symbols = len(symbols)
        self.ev = ev
        rows, cols, vals = [], [], []
        self.dynamic: list = []  # (slot, symbol index, expr, symbol)
        for slot, e in enumerate(slot_exprs):
            for sym in sorted(getattr(e, "free_symbols", ()), key=lambda x: x.name):
                if sym not in index:
                    continue
                d = sympy.diff(e, sym)
                if not d.free_symbols:
                    rows.append(slot)
                    cols.append(index[sym])
                    vals.append(float(d))
                else:
                    self.dynamic.append((slot, index[sym], e, sym))
        self.rows = np.array(rows, dtype=np.int64)
        self.cols = np.array(cols, dtype=np.int64)
        self.vals = np.array(vals, dtype=np.float64)

    def __call__(self, gs: np.ndarray, symvals: np.ndarray) -> np.ndarray:
        """gs: (B, n_slots); symvals: (n_symbols, B).  Returns (B, n_symbols)."""
        gs = np.asarray(gs, dtype=np.float64)
        out = np.zeros((gs.shape[0], self.n_symbols))
        if self.rows.size:
            np.add.at(out.T, self.cols, (gs[:, self.rows] * self.vals).T)
        for slot, i, e, sym in self.dynamic:
            d = self.ev.dvalues(e, sym, symvals)
            if d is not None:
                out[:, i] += gs[:, slot] * d
        return out


def _batch_group(gate):
    """Gates whose matrix batch depends only on their parameter values and this key can be
    evaluated together: one _matrix_batch call over every gate of the group."""
    from .gates import EigenGate, FSimGate, PhasedXPowGate, PhasedXZGate

    if isinstance(gate, (EigenGate, PhasedXPowGate)):
        return type(gate), float(gate.global_shift)
    if isinstance(gate, (PhasedXZGate, FSimGate)):
        return type(gate), 0.0
    return None


class _ExprBatch:
    """Values of a list of parameter expressions at B points, as one (len, B) array: affine
    expressions (a * symbol + b, the usual case) in one vectorised step, the rest one by one."""

    def __init__(self, exprs: Sequence, ev: ParamEvaluator):
        self.n = len(exprs)
        comp = []
        for e in exprs:
            c = ev._fns.get(e)
            if c is None:
                c = ev._fns[e] = ev._compile(e)
            comp.append(c)
        self.aff = np.array([i for i, c in enumerate(comp) if c[0] == "affine"], dtype=np.int64)
        self.aff_sym = np.array([comp[i][1][0] for i in self.aff], dtype=np.int64)
        self.aff_a = np.array([comp[i][1][1] for i in self.aff], dtype=np.float64)
        self.aff_b = np.array([comp[i][1][2] for i in self.aff], dtype=np.float64)
        self.const = np.array([i for i, c in enumerate(comp) if c[0] == "const"], dtype=np.int64)
        self.const_v = np.array([comp[i][1] for i in self.const], dtype=np.float64)
        self.other = [(i, c) for i, c in enumerate(comp) if c[0] not in ("affine", "const")]

    def __call__(self, symvals: np.ndarray) -> np.ndarray:
        out = np.empty((self.n, symvals.shape[1]))
        if self.aff.size:
            out[self.aff] = self.aff_a[:, None] * symvals[self.aff_sym] + self.aff_b[:, None]
        if self.const.size:
            out[self.const] = self.const_v[:, None]
        for i, c in self.other:
            out[i] = ParamEvaluator._eval(c, symvals)
        return out


class ParamPlan:
    """`fill_param_mats` compiled once per (program, evaluator).

    Every parameterised entry of a program's matrix buffer is a transform of one gate's matrix
    (or of its derivative): the lowering's payload transforms (conjugate transposes, target
    permutations, control embeddings, diagonal extraction, derivative payloads) are affine in the
    gate matrix's real and imaginary parts.  Each entry's affine map is read off its transform
    once, from the images of 0 and of the 2 d^2 real unit matrices, and checked on a random
    matrix; the maps of all affine entries form one sparse matrix, so filling the buffer is one
    batched matrix evaluation per gate family, one sparse product and one scatter.  Entries whose
    transform is not affine (the FP16 split of an MFMA block payload, density-matrix diagonals
    d (x) conj(d)) keep the per-entry call."""

    def __init__(self, prog: PackedProgram, ev: ParamEvaluator):
        import scipy.sparse

        self.base = prog.mats_arr
        refs: dict = {}
        for _off, ref, _size in prog.param_entries:
            refs.setdefault((id(ref.gate), ref.deriv), ref)
        grouped: dict = {}
        for key, ref in refs.items():
            g = _batch_group(ref.gate)
            grouped.setdefault(key if g is None else (g, ref.deriv), []).append(key)
        # groups: (representative ref, keys, one _ExprBatch per parameter position)
        self.groups = []
        where: dict = {}  # key -> (group index, position in group)
        for keys in grouped.values():
            rep = refs[keys[0]]
            exprs = [_ExprBatch([refs[k].exprs[p] for k in keys], ev) for p in range(len(rep.exprs))]
            for j, k in enumerate(keys):
                where[k] = (len(self.groups), j)
            self.groups.append((rep, keys, exprs))
        # input layout: group blocks of (G, d, d) complex, as float64 pairs
        zero = np.zeros((len(ev.symbols), 1))
        self.in_off, n_in = [], 0
        dims = []
        for rep, keys, exprs in self.groups:
            mb = self._group_matrices(rep, exprs, zero)
            d = mb.shape[-1]
            dims.append(d)
            self.in_off.append(n_in)
            n_in += len(keys) * 2 * d * d
        self.n_in = n_in
        rng = np.random.default_rng(0)
        rows, cols, vals, crow, cval = [], [], [], [], []
        self.nonaffine = []
        for off, ref, size in prog.param_entries:
            gi, j = where[(id(ref.gate), ref.deriv)]
            d = dims[gi]
            nq = 2 * d * d
            if d > 8:  # wide gates: not worth 2 d^2 probe calls
                self.nonaffine.append((off, ref, size))
                continue
            probe = np.zeros((1 + nq, d, d), dtype=np.complex128)
            for q in range(nq):
                probe[1 + q].reshape(-1)[q // 2] = 1.0 if q % 2 == 0 else 1.0j
            img = np.asarray(ref.transform(probe), dtype=np.complex128).reshape(1 + nq, -1)
            c, A = img[0], img[1:] - img[0]
            y = rng.standard_normal((2, d, d)) + 1j * rng.standard_normal((2, d, d))
            want = np.asarray(ref.transform(y), dtype=np.complex128).reshape(2, -1)
            yr = y.reshape(2, -1).view(np.float64)
            got = c + yr @ A
            if img.shape[1] != size or not np.allclose(got, want, rtol=0, atol=1e-12 * (1 + np.abs(want).max())):
                self.nonaffine.append((off, ref, size))
                continue
            base_col = self.in_off[gi] + j * nq
            for part, M in ((0, A.real), (1, A.imag)):
                q, m = np.nonzero(M)
                rows.append(2 * (off + m) + part)
                cols.append(base_col + q)
                vals.append(M[q, m])
            crow.append(2 * (off + np.arange(size)))
            cval.append(c.real)
            crow.append(2 * (off + np.arange(size)) + 1)
            cval.append(c.imag)
        self.dst = np.concatenate(crow) if crow else np.zeros(0, dtype=np.int64)  # float32 positions written
        self.const = np.concatenate(cval) if cval else np.zeros(0)
        pos = np.full(2 * self.base.size, -1, dtype=np.int64)
        pos[self.dst] = np.arange(self.dst.size)
        r = np.concatenate(rows) if rows else np.zeros(0, dtype=np.int64)
        self.S = scipy.sparse.csr_matrix(
            (np.concatenate(vals) if vals else np.zeros(0), (pos[r], np.concatenate(cols) if cols else r)),
            shape=(self.dst.size, n_in),
        )

    @staticmethod
    def _group_matrices(rep, exprs, symvals):
        """(G * B, d, d) matrices (or derivatives) of the group's gates, gate-major."""
        vals = tuple(e(symvals).reshape(-1) for e in exprs)
        if rep.deriv is None:
            return np.asarray(rep.gate._matrix_batch(vals))
        return np.asarray(rep.gate._dmatrix_batch(vals, rep.deriv))

    def __call__(self, symvals: np.ndarray) -> np.ndarray:
        B = symvals.shape[1]
        out = np.broadcast_to(self.base, (B, self.base.size)).copy()
        x = np.empty((B, self.n_in))
        mats = {}
        for (rep, keys, exprs), off in zip(self.groups, self.in_off):
            mb = self._group_matrices(rep, exprs, symvals)
            d = mb.shape[-1]
            mb = mb.reshape(len(keys), B, d, d)
            x[:, off : off + len(keys) * 2 * d * d] = (
                np.ascontiguousarray(mb.transpose(1, 0, 2, 3)).reshape(B, -1).view(np.float64))
            if self.nonaffine:
                for j, k in enumerate(keys):
                    mats[k] = mb[j]
        if self.dst.size:
            y = (self.S @ x.T).T + self.const
            out.view(np.float32)[:, self.dst] = y
        for off, ref, size in self.nonaffine:
            payload = ref.transform(mats[(id(ref.gate), ref.deriv)])
            out[:, off : off + size] = payload.reshape(B, -1)
        return out


def fill_param_mats(prog: PackedProgram, ev: ParamEvaluator, symvals: np.ndarray) -> np.ndarray:
    """(B, mats) complex64: the program's matrix buffer for each of B resolvers (through the
    program's ParamPlan for this evaluator, built on first use)."""
    plan = getattr(prog, "_param_plan", None)
    if plan is None or plan[0] is not ev:
        plan = (ev, ParamPlan(prog, ev))
        prog._param_plan = plan
    return plan[1](symvals)


def fill_param_mats_direct(prog: PackedProgram, ev: ParamEvaluator, symvals: np.ndarray) -> np.ndarray:
    """fill_param_mats by calling every entry's transform (the definition ParamPlan compiles)."""
    B = symvals.shape[1]
    out = np.broadcast_to(prog.mats_arr, (B, prog.mats_arr.size)).copy()
    refs: dict = {}
    for _off, ref, _size in prog.param_entries:
        refs.setdefault((id(ref.gate), ref.deriv), ref)
    groups: dict = {}
    for key, ref in refs.items():
        g = _batch_group(ref.gate)
        groups.setdefault(None if g is None else (g, ref.deriv), []).append(key)
    cache: dict = {}
    for gk, keys in groups.items():
        if gk is None:
            for key in keys:
                ref = refs[key]
                vals = tuple(ev.values(e, symvals) for e in ref.exprs)
                cache[key] = (ref.gate._matrix_batch(vals) if ref.deriv is None
                              else ref.gate._dmatrix_batch(vals, ref.deriv))
            continue
        rep = refs[keys[0]]
        G = len(keys)
        vals = tuple(
            np.concatenate([ev.values(refs[key].exprs[p], symvals) for key in keys]) for p in range(len(rep.exprs))
        )
        mb = rep.gate._matrix_batch(vals) if rep.deriv is None else rep.gate._dmatrix_batch(vals, rep.deriv)
        mb = mb.reshape((G, B) + mb.shape[1:])
        for i, key in enumerate(keys):
            cache[key] = mb[i]
    for off, ref, size in prog.param_entries:
        payload = ref.transform(cache[(id(ref.gate), ref.deriv)])
        out[:, off : off + size] = payload.reshape(B, -1)
    return out