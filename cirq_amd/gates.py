"""Gate and operation types with Cirq's semantics (matrices, global phases, parameters).

Every unitary here follows the reference definition it cites; matrices are complex128 and include
Cirq's exact global phase, because fused blocks must equal the product of the original gates
(cirq-core/cirq/transformers/merge_k_qubit_gates.py:55-63 multiplies unitaries exactly).

Parameterised gates expose ``_param_exprs()`` (numbers or sympy expressions) and
``_matrix_batch(vals)`` which evaluates the matrix for a batch of parameter values at once
(shape (B, d, d)); the simulator resolves a whole sweep with one call per gate.
"""

from __future__ import annotations

import cmath
import math
import numbers
from typing import Any, Iterable, Sequence

import numpy as np
import sympy


def is_parameterized(x) -> bool:
    """True if x is a sympy expression with free symbols or an object reporting parameters."""
    if isinstance(x, sympy.Basic):
        return bool(x.free_symbols)
    f = getattr(x, "_is_parameterized_", None)
    if f is not None:
        r = f()
        return bool(r) if r is not NotImplemented else False
    return False


def _pi(x):
    return sympy.pi if isinstance(x, sympy.Basic) else np.pi


def _canon(x):
    """float for numbers, unchanged sympy expression otherwise."""
    if isinstance(x, sympy.Basic):
        if x.free_symbols:
            return x
        return float(x)
    return float(x)


def _phase_factor(e, h):
    """Cirq's ``1j ** (2 * e * h)`` for a float exponent (cirq-core/cirq/ops/eigen_gate.py:300-304)."""
    return 1j ** (2 * e * h)


# --------------------------------------------------------------------------------------------
class Gate:
    """Base gate.  Subclasses define ``_num_qubits_`` and either ``_unitary_`` or channels."""

    #: True when the matrix is diagonal for every parameter value (lowered to a phase op).
    _diagonal = False

    def num_qubits(self) -> int:
        return self._num_qubits_()

    def _num_qubits_(self) -> int:
        raise NotImplementedError

    def _qid_shape_(self):
        return (2,) * self.num_qubits()

    def on(self, *qubits) -> "GateOperation":
        return GateOperation(self, qubits)

    def __call__(self, *qubits):
        return self.on(*qubits)

    def on_each(self, *targets) -> list["GateOperation"]:
        out = []
        for t in targets:
            if isinstance(t, Iterable) and not hasattr(t, "dimension"):
                out.extend(self.on_each(*t))
            else:
                out.append(self.on(t))
        return out

    def _has_unitary_(self) -> bool:
        """True when the gate defines its own matrix (a subclass that only decomposes does not)."""
        cls = type(self)
        own = cls._matrix_batch is not Gate._matrix_batch or cls._unitary_ is not Gate._unitary_
        return own and not self._is_parameterized_()

    def _is_parameterized_(self) -> bool:
        return any(is_parameterized(v) for v in self._param_exprs())

    def _parameter_names_(self) -> set:
        names = set()
        for v in self._param_exprs():
            if isinstance(v, sympy.Basic):
                names |= {s.name for s in v.free_symbols}
        return names

    def _param_exprs(self) -> tuple:
        return ()

    def _unitary_(self):
        if self._is_parameterized_():
            return NotImplemented
        vals = tuple(np.array([float(v)]) for v in self._param_exprs())
        return self._matrix_batch(vals)[0]

    def _matrix_batch(self, vals: tuple) -> np.ndarray:
        raise NotImplementedError

    def _dmatrix_batch(self, vals: tuple, k: int) -> np.ndarray:
        """d(matrix)/d(parameter k), by a central difference of step 1e-6 (error ~1e-10)."""
        h = 1e-6
        up = tuple(v + h if i == k else v for i, v in enumerate(vals))
        dn = tuple(v - h if i == k else v for i, v in enumerate(vals))
        return (self._matrix_batch(up) - self._matrix_batch(dn)) / (2 * h)

    def _with_params(self, vals: Sequence) -> "Gate":
        raise NotImplementedError

    def _resolve_parameters_(self, resolver, recursive: bool = True) -> "Gate":
        if not self._is_parameterized_():
            return self
        vals = [resolver.value_of(v, recursive) for v in self._param_exprs()]
        return self._with_params([_canon(v) for v in vals])

    def controlled(self, num_controls: int | None = None, control_values=None) -> "ControlledGate":
        if num_controls is None:
            num_controls = 1 if control_values is None else len(control_values)
        return ControlledGate(self, num_controls=num_controls, control_values=control_values)

    def __pow__(self, power):
        return NotImplemented


# --------------------------------------------------------------------------------------------
class EigenGate(Gate):
    """U(t) = sum_k P_k * exp(i pi t (h_k + s)) (cirq-core/cirq/ops/eigen_gate.py:295-305)."""

    def __init__(self, *, exponent=1.0, global_shift: float = 0.0):
        self._exponent = _canon(exponent) if not isinstance(exponent, sympy.Basic) else exponent
        self._global_shift = float(global_shift)

    @property
    def exponent(self):
        return self._exponent

    @property
    def global_shift(self) -> float:
        return self._global_shift

    def _eigen_components(self) -> list:
        raise NotImplementedError

    def _num_qubits_(self) -> int:
        return int(round(math.log2(self._eigen_components()[0][1].shape[0])))

    def _param_exprs(self) -> tuple:
        return (self._exponent,)

    def _unitary_(self):
        if self._is_parameterized_():
            return NotImplemented
        e = float(self._exponent)
        return np.sum(
            [comp * _phase_factor(e, h + self._global_shift) for h, comp in self._eigen_components()],
            axis=0,
        ).astype(np.complex128)

    def _matrix_batch(self, vals):
        (t,) = vals
        t = np.asarray(t, dtype=np.float64)
        out = 0
        for h, comp in self._eigen_components():
            ph = np.exp(1j * np.pi * t * (h + self._global_shift))
            out = out + ph[:, None, None] * np.asarray(comp, dtype=np.complex128)[None]
        return out

    def _dmatrix_batch(self, vals, k):
        (t,) = vals
        t = np.asarray(t, dtype=np.float64)
        out = 0
        for h, comp in self._eigen_components():
            w = 1j * np.pi * (h + self._global_shift)
            ph = w * np.exp(1j * np.pi * t * (h + self._global_shift))
            out = out + ph[:, None, None] * np.asarray(comp, dtype=np.complex128)[None]
        return out

    def _with_exponent(self, exponent) -> "EigenGate":
        return type(self)(exponent=exponent, global_shift=self._global_shift)

    def _with_params(self, vals):
        return self._with_exponent(vals[0])

    def __pow__(self, power):
        if not isinstance(power, (numbers.Number, sympy.Basic)):
            return NotImplemented
        return self._with_exponent(self._exponent * power)

    def _value_key(self):
        return (type(self).__name__, self._exponent, self._global_shift)

    def __eq__(self, other):
        if not isinstance(other, EigenGate):
            return NotImplemented
        return self._value_key() == other._value_key()

    def __hash__(self):
        return hash(self._value_key())

    def _fmt(self, name):
        if self._global_shift == 0:
            return name if self._exponent == 1 else f"{name}**{self._exponent}"
        return f"{type(self).__name__}(exponent={self._exponent}, global_shift={self._global_shift})"


# ---- single-qubit eigen gates (cirq-core/cirq/ops/common_gates.py) --------------------------
_X_COMPS = [
    (0, np.array([[0.5, 0.5], [0.5, 0.5]])),
    (1, np.array([[0.5, -0.5], [-0.5, 0.5]])),
]
_Y_COMPS = [
    (0, np.array([[0.5, -0.5j], [0.5j, 0.5]])),
    (1, np.array([[0.5, 0.5j], [-0.5j, 0.5]])),
]
_Z_COMPS = [(0, np.diag([1.0, 0.0])), (1, np.diag([0.0, 1.0]))]
_S2 = np.sqrt(2)
_H_COMPS = [
    (0, np.array([[3 + 2 * _S2, 1 + _S2], [1 + _S2, 1]]) / (4 + 2 * _S2)),
    (1, np.array([[3 - 2 * _S2, 1 - _S2], [1 - _S2, 1]]) / (4 - 2 * _S2)),
]


class XPowGate(EigenGate):
    """cirq.XPowGate (cirq-core/cirq/ops/common_gates.py:63-160)."""

    def _eigen_components(self):
        return _X_COMPS

    def __repr__(self):
        return self._fmt("cirq.X")


class YPowGate(EigenGate):
    """cirq.YPowGate (cirq-core/cirq/ops/common_gates.py:358-478)."""

    def _eigen_components(self):
        return _Y_COMPS

    def __repr__(self):
        return self._fmt("cirq.Y")


class ZPowGate(EigenGate):
    """cirq.ZPowGate (cirq-core/cirq/ops/common_gates.py:598-765)."""

    _diagonal = True

    def _eigen_components(self):
        return _Z_COMPS

    def __repr__(self):
        return self._fmt("cirq.Z")


class HPowGate(EigenGate):
    """cirq.HPowGate (cirq-core/cirq/ops/common_gates.py:924-960)."""

    def _eigen_components(self):
        return _H_COMPS

    def __repr__(self):
        return self._fmt("cirq.H")


class _RotationMixin:
    """Rx/Ry/Rz: exponent = rads / pi, global_shift = -1/2 (cirq-core/cirq/ops/common_gates.py:304-330)."""

    _base = None

    def __init__(self, *, rads):
        self._rads = rads if isinstance(rads, sympy.Basic) else float(rads)
        super().__init__(exponent=self._rads / _pi(self._rads), global_shift=-0.5)

    def _with_exponent(self, exponent):
        return type(self)(rads=exponent * _pi(exponent))

    @property
    def rads(self):
        return self._rads

    def __repr__(self):
        return f"cirq.{type(self).__name__.lower()}({self._rads!r})"


class Rx(_RotationMixin, XPowGate):
    pass


class Ry(_RotationMixin, YPowGate):
    pass


class Rz(_RotationMixin, ZPowGate):
    pass


def rx(rads) -> Rx:
    return Rx(rads=rads)


def ry(rads) -> Ry:
    return Ry(rads=rads)


def rz(rads) -> Rz:
    return Rz(rads=rads)


# ---- two- and three-qubit eigen gates --------------------------------------------------------
class CZPowGate(EigenGate):
    """cirq.CZPowGate (cirq-core/cirq/ops/common_gates.py:1043-1082)."""

    _diagonal = True

    def _eigen_components(self):
        return [(0, np.diag([1.0, 1, 1, 0])), (1, np.diag([0.0, 0, 0, 1]))]

    def __repr__(self):
        return self._fmt("cirq.CZ")


class CXPowGate(EigenGate):
    """cirq.CXPowGate (cirq-core/cirq/ops/common_gates.py:1199-1260)."""

    def _eigen_components(self):
        return [
            (0, np.array([[1.0, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0.5, 0.5], [0, 0, 0.5, 0.5]])),
            (1, np.array([[0.0, 0, 0, 0], [0, 0, 0, 0], [0, 0, 0.5, -0.5], [0, 0, -0.5, 0.5]])),
        ]

    def __repr__(self):
        return self._fmt("cirq.CNOT")


CNotPowGate = CXPowGate


class SwapPowGate(EigenGate):
    """cirq.SwapPowGate (cirq-core/cirq/ops/swap_gates.py:86-97)."""

    def _eigen_components(self):
        return [
            (0, np.array([[1.0, 0, 0, 0], [0, 0.5, 0.5, 0], [0, 0.5, 0.5, 0], [0, 0, 0, 1]])),
            (1, np.array([[0.0, 0, 0, 0], [0, 0.5, -0.5, 0], [0, -0.5, 0.5, 0], [0, 0, 0, 0]])),
        ]

    def __repr__(self):
        return self._fmt("cirq.SWAP")


class ISwapPowGate(EigenGate):
    """cirq.ISwapPowGate (cirq-core/cirq/ops/swap_gates.py:207-219)."""

    def _eigen_components(self):
        return [
            (0, np.diag([1.0, 0, 0, 1])),
            (0.5, np.array([[0.0, 0, 0, 0], [0, 0.5, 0.5, 0], [0, 0.5, 0.5, 0], [0, 0, 0, 0]])),
            (-0.5, np.array([[0.0, 0, 0, 0], [0, 0.5, -0.5, 0], [0, -0.5, 0.5, 0], [0, 0, 0, 0]])),
        ]

    def __repr__(self):
        return self._fmt("cirq.ISWAP")


class XXPowGate(EigenGate):
    """cirq.XXPowGate (cirq-core/cirq/ops/parity_gates.py:69-81)."""

    def _eigen_components(self):
        return [
            (0.0, np.array([[0.5, 0, 0, 0.5], [0, 0.5, 0.5, 0], [0, 0.5, 0.5, 0], [0.5, 0, 0, 0.5]])),
            (1.0, np.array([[0.5, 0, 0, -0.5], [0, 0.5, -0.5, 0], [0, -0.5, 0.5, 0], [-0.5, 0, 0, 0.5]])),
        ]

    def __repr__(self):
        return self._fmt("cirq.XX")


class YYPowGate(EigenGate):
    """cirq.YYPowGate (cirq-core/cirq/ops/parity_gates.py:155-169)."""

    def _eigen_components(self):
        return [
            (0.0, np.array([[0.5, 0, 0, -0.5], [0, 0.5, 0.5, 0], [0, 0.5, 0.5, 0], [-0.5, 0, 0, 0.5]])),
            (1.0, np.array([[0.5, 0, 0, 0.5], [0, 0.5, -0.5, 0], [0, -0.5, 0.5, 0], [0.5, 0, 0, 0.5]])),
        ]

    def __repr__(self):
        return self._fmt("cirq.YY")


class ZZPowGate(EigenGate):
    """cirq.ZZPowGate (cirq-core/cirq/ops/parity_gates.py:237-241)."""

    _diagonal = True

    def _eigen_components(self):
        return [(0, np.diag([1.0, 0, 0, 1])), (1, np.diag([0.0, 1, 1, 0]))]

    def __repr__(self):
        return self._fmt("cirq.ZZ")


class CCZPowGate(EigenGate):
    """cirq.CCZPowGate (cirq-core/cirq/ops/three_qubit_gates.py:63-64)."""

    _diagonal = True

    def _eigen_components(self):
        return [(0, np.diag([1.0, 1, 1, 1, 1, 1, 1, 0])), (1, np.diag([0.0, 0, 0, 0, 0, 0, 0, 1]))]

    def __repr__(self):
        return self._fmt("cirq.CCZ")


class CCXPowGate(EigenGate):
    """cirq.CCXPowGate (cirq-core/cirq/ops/three_qubit_gates.py:394-403)."""

    def _eigen_components(self):
        c0 = np.zeros((8, 8))
        c0[:6, :6] = np.eye(6)
        c0[6:, 6:] = [[0.5, 0.5], [0.5, 0.5]]
        c1 = np.zeros((8, 8))
        c1[6:, 6:] = [[0.5, -0.5], [-0.5, 0.5]]
        return [(0, c0), (1, c1)]

    def __repr__(self):
        return self._fmt("cirq.TOFFOLI")


class CSwapGate(Gate):
    """cirq.CSwapGate (Fredkin): swaps targets when the control is 1."""

    def _num_qubits_(self):
        return 3

    def _param_exprs(self):
        return ()

    def _matrix_batch(self, vals):
        u = np.eye(8, dtype=np.complex128)
        u[[5, 6]] = u[[6, 5]]
        return u[None]

    def _unitary_(self):
        return self._matrix_batch(())[0]

    def _with_params(self, vals):
        return self

    def __repr__(self):
        return "cirq.FREDKIN"


# ---- gates with several parameters -----------------------------------------------------------
class PhasedXPowGate(Gate):
    """Z^p X^t Z^-p with phase exp(i pi s t) (cirq-core/cirq/ops/phased_x_gate.py:138-145)."""

    def __init__(self, *, phase_exponent, exponent=1.0, global_shift: float = 0.0):
        self._phase_exponent = phase_exponent if isinstance(phase_exponent, sympy.Basic) else float(phase_exponent)
        self._exponent = exponent if isinstance(exponent, sympy.Basic) else float(exponent)
        self._global_shift = float(global_shift)

    @property
    def phase_exponent(self):
        return self._phase_exponent

    @property
    def exponent(self):
        return self._exponent

    @property
    def global_shift(self):
        return self._global_shift

    def _num_qubits_(self):
        return 1

    def _param_exprs(self):
        return (self._phase_exponent, self._exponent)

    def _unitary_(self):
        if self._is_parameterized_():
            return NotImplemented
        z = ZPowGate(exponent=self._phase_exponent)._unitary_()
        x = XPowGate(exponent=self._exponent)._unitary_()
        p = np.exp(1j * np.pi * self._global_shift * self._exponent)
        return np.dot(np.dot(z, x), np.conj(z)) * p

    def _matrix_batch(self, vals):
        p, t = (np.asarray(v, dtype=np.float64) for v in vals)
        p, t = np.broadcast_arrays(p, t)
        z = np.zeros(p.shape + (2, 2), dtype=np.complex128)
        z[..., 0, 0] = 1
        z[..., 1, 1] = np.exp(1j * np.pi * p)
        x = XPowGate(exponent=0)._matrix_batch((t,))
        ph = np.exp(1j * np.pi * self._global_shift * t)
        return z @ x @ np.conj(z) * ph[:, None, None]

    def _with_params(self, vals):
        return PhasedXPowGate(phase_exponent=vals[0], exponent=vals[1], global_shift=self._global_shift)

    def __pow__(self, power):
        return PhasedXPowGate(
            phase_exponent=self._phase_exponent, exponent=self._exponent * power, global_shift=self._global_shift
        )

    def __eq__(self, other):
        if not isinstance(other, PhasedXPowGate):
            return NotImplemented
        return (self._phase_exponent, self._exponent, self._global_shift) == (
            other._phase_exponent,
            other._exponent,
            other._global_shift,
        )

    def __hash__(self):
        return hash(("PhasedXPowGate", self._phase_exponent, self._exponent, self._global_shift))

    def __repr__(self):
        return f"cirq.PhasedXPowGate(phase_exponent={self._phase_exponent!r}, exponent={self._exponent!r})"


class PhasedXZGate(Gate):
    """cirq.PhasedXZGate (cirq-core/cirq/ops/phased_x_z_gate.py:187-207)."""

    def __init__(self, *, x_exponent, z_exponent, axis_phase_exponent):
        f = lambda v: v if isinstance(v, sympy.Basic) else float(v)
        self._x_exponent = f(x_exponent)
        self._z_exponent = f(z_exponent)
        self._axis_phase_exponent = f(axis_phase_exponent)

    x_exponent = property(lambda self: self._x_exponent)
    z_exponent = property(lambda self: self._z_exponent)
    axis_phase_exponent = property(lambda self: self._axis_phase_exponent)

    def _num_qubits_(self):
        return 1

    def _param_exprs(self):
        return (self._x_exponent, self._z_exponent, self._axis_phase_exponent)

    def _unitary_(self):
        if self._is_parameterized_():
            return NotImplemented
        a, x, z = self._axis_phase_exponent, self._x_exponent, self._z_exponent
        if x % 1 == 0.5:
            ucx = 0.5 + 0.5j * (-1) ** math.floor(x)
            usx = ucx.conjugate()
        else:
            cpxh = 1j**x
            ucx = cpxh * cpxh.real
            usx = -1j * cpxh * cpxh.imag
        return np.array(
            [[ucx, usx * 1j ** (-2 * a)], [usx * 1j ** (2 * (z + a)), ucx * 1j ** (2 * z)]], dtype=np.complex128
        )

    def _matrix_batch(self, vals):
        x, z, a = np.broadcast_arrays(*(np.asarray(v, dtype=np.float64) for v in vals))
        cpxh = np.exp(0.5j * np.pi * x)
        ucx = cpxh * cpxh.real
        usx = -1j * cpxh * cpxh.imag
        u = np.empty(x.shape + (2, 2), dtype=np.complex128)
        u[..., 0, 0] = ucx
        u[..., 0, 1] = usx * np.exp(-1j * np.pi * a)
        u[..., 1, 0] = usx * np.exp(1j * np.pi * (z + a))
        u[..., 1, 1] = ucx * np.exp(1j * np.pi * z)
        return u

    def _with_params(self, vals):
        return PhasedXZGate(x_exponent=vals[0], z_exponent=vals[1], axis_phase_exponent=vals[2])

    def __repr__(self):
        return (
            f"cirq.PhasedXZGate(x_exponent={self._x_exponent!r}, z_exponent={self._z_exponent!r}, "
            f"axis_phase_exponent={self._axis_phase_exponent!r})"
        )


class FSimGate(Gate):
    """cirq.FSimGate (cirq-core/cirq/ops/fsim_gate.py:135-150)."""

    def __init__(self, theta, phi):
        f = lambda v: v if isinstance(v, sympy.Basic) else float(v)
        self._theta = f(theta)
        self._phi = f(phi)

    theta = property(lambda self: self._theta)
    phi = property(lambda self: self._phi)

    def _num_qubits_(self):
        return 2

    def _param_exprs(self):
        return (self._theta, self._phi)

    def _unitary_(self):
        if self._is_parameterized_():
            return NotImplemented
        a = math.cos(self._theta)
        b = -1j * math.sin(self._theta)
        c = cmath.exp(-1j * self._phi)
        return np.array([[1, 0, 0, 0], [0, a, b, 0], [0, b, a, 0], [0, 0, 0, c]], dtype=np.complex128)

    def _matrix_batch(self, vals):
        th, ph = np.broadcast_arrays(*(np.asarray(v, dtype=np.float64) for v in vals))
        u = np.zeros(th.shape + (4, 4), dtype=np.complex128)
        u[..., 0, 0] = 1
        u[..., 1, 1] = u[..., 2, 2] = np.cos(th)
        u[..., 1, 2] = u[..., 2, 1] = -1j * np.sin(th)
        u[..., 3, 3] = np.exp(-1j * ph)
        return u

    def _with_params(self, vals):
        return FSimGate(vals[0], vals[1])

    def __repr__(self):
        return f"cirq.FSimGate(theta={self._theta!r}, phi={self._phi!r})"


class MatrixGate(Gate):
    """A fixed unitary (cirq-core/cirq/ops/matrix_gates.py:42-236)."""

    def __init__(self, matrix, *, name: str | None = None, qid_shape=None):
        m = np.asarray(matrix, dtype=np.complex128)
        if m.ndim != 2 or m.shape[0] != m.shape[1] or m.shape[0] & (m.shape[0] - 1):
            raise ValueError("MatrixGate needs a square 2^k x 2^k matrix")
        self._matrix = m
        self._name = name

    def _num_qubits_(self):
        return int(round(math.log2(self._matrix.shape[0])))

    def _unitary_(self):
        return np.array(self._matrix)

    def _matrix_batch(self, vals):
        return self._matrix[None]

    def _with_params(self, vals):
        return self

    def _has_unitary_(self):
        return True

    def __pow__(self, power):
        if power == -1:
            return MatrixGate(self._matrix.conj().T)
        w, v = np.linalg.eig(self._matrix)
        return MatrixGate(v @ np.diag(w.astype(np.complex128) ** power) @ np.linalg.inv(v))

    def __repr__(self):
        return f"cirq.MatrixGate({self._matrix!r})"


class DiagonalGate(Gate):
    """diag(exp(i*angle_k)) (cirq-core/cirq/ops/diagonal_gate.py:100-122)."""

    _diagonal = True

    def __init__(self, diag_angles_radians: Sequence):
        self._angles = tuple(a if isinstance(a, sympy.Basic) else float(a) for a in diag_angles_radians)
        if len(self._angles) & (len(self._angles) - 1) or not self._angles:
            raise ValueError("DiagonalGate needs 2^k angles")

    @property
    def diag_angles_radians(self):
        return self._angles

    def _num_qubits_(self):
        return int(round(math.log2(len(self._angles))))

    def _param_exprs(self):
        return self._angles

    def _matrix_batch(self, vals):
        ang = np.stack(np.broadcast_arrays(*(np.asarray(v, dtype=np.float64) for v in vals)), axis=-1)
        d = np.exp(1j * ang)
        out = np.zeros(d.shape + (d.shape[-1],), dtype=np.complex128)
        idx = np.arange(d.shape[-1])
        out[..., idx, idx] = d
        return out

    def _unitary_(self):
        if self._is_parameterized_():
            return NotImplemented
        return np.diag([np.exp(1j * a) for a in self._angles])

    def _with_params(self, vals):
        return DiagonalGate(vals)

    def __repr__(self):
        return f"cirq.DiagonalGate({list(self._angles)!r})"


class IdentityGate(Gate):
    """cirq.IdentityGate."""

    _diagonal = True

    def __init__(self, num_qubits: int = 1):
        self._n = int(num_qubits)

    def _num_qubits_(self):
        return self._n

    def _matrix_batch(self, vals):
        return np.eye(2**self._n, dtype=np.complex128)[None]

    def _unitary_(self):
        return np.eye(2**self._n, dtype=np.complex128)

    def _with_params(self, vals):
        return self

    def __pow__(self, power):
        return self

    def __repr__(self):
        return "cirq.I" if self._n == 1 else f"cirq.IdentityGate({self._n})"


class GlobalPhaseGate(Gate):
    """A zero-qubit gate multiplying the state by ``coefficient`` (cirq.GlobalPhaseGate)."""

    _diagonal = True

    def __init__(self, coefficient):
        self._coefficient = coefficient if isinstance(coefficient, sympy.Basic) else complex(coefficient)

    @property
    def coefficient(self):
        return self._coefficient

    def _num_qubits_(self):
        return 0

    def _param_exprs(self):
        return ()

    def _matrix_batch(self, vals):
        return np.array([[[self._coefficient]]], dtype=np.complex128)

    def _unitary_(self):
        return np.array([[self._coefficient]], dtype=np.complex128)

    def _with_params(self, vals):
        return self

    def __repr__(self):
        return f"cirq.GlobalPhaseGate({self._coefficient!r})"


def global_phase_operation(coefficient) -> "GateOperation":
    return GateOperation(GlobalPhaseGate(coefficient), ())


class ControlledGate(Gate):
    """sub_gate applied when the control qubits (first) hold control_values (cirq.ControlledGate)."""

    def __init__(self, sub_gate: Gate, num_controls: int = 1, control_values=None):
        self._sub_gate = sub_gate
        self._num_controls = int(num_controls)
        if control_values is None:
            control_values = (1,) * self._num_controls
        cv = []
        for v in control_values:
            if isinstance(v, (tuple, list)):
                if len(v) != 1:
                    raise ValueError("only single control values per qubit are supported")
                v = v[0]
            cv.append(int(v))
        self._control_values = tuple(cv)

    sub_gate = property(lambda self: self._sub_gate)
    num_controls = property(lambda self: self._num_controls)
    control_values = property(lambda self: self._control_values)

    def _num_qubits_(self):
        return self._num_controls + self._sub_gate.num_qubits()

    @property
    def _diagonal(self):
        return getattr(self._sub_gate, "_diagonal", False)

    def _param_exprs(self):
        return self._sub_gate._param_exprs()

    def _embed(self, sub):
        nc = self._num_controls
        ds = sub.shape[-1]
        dim = (2**nc) * ds
        out = np.broadcast_to(np.eye(dim, dtype=np.complex128), sub.shape[:-2] + (dim, dim)).copy()
        ctl = 0
        for v in self._control_values:
            ctl = ctl * 2 + v
        out[..., ctl * ds : (ctl + 1) * ds, ctl * ds : (ctl + 1) * ds] = sub
        return out

    def _matrix_batch(self, vals):
        return self._embed(self._sub_gate._matrix_batch(vals))

    def _dmatrix_batch(self, vals, k):
        d = self._sub_gate._dmatrix_batch(vals, k)
        e = self._embed(d)
        # the identity blocks do not depend on the parameter
        nc = self._num_controls
        ds = d.shape[-1]
        ctl = 0
        for v in self._control_values:
            ctl = ctl * 2 + v
        mask = np.zeros(e.shape[-2:], dtype=bool)
        mask[ctl * ds : (ctl + 1) * ds, ctl * ds : (ctl + 1) * ds] = True
        return np.where(mask, e, 0)

    def _unitary_(self):
        u = self._sub_gate._unitary_()
        if u is NotImplemented or u is None:
            return NotImplemented
        return self._embed(np.asarray(u)[None])[0]

    def _with_params(self, vals):
        return ControlledGate(self._sub_gate._with_params(vals), self._num_controls, self._control_values)

    def _resolve_parameters_(self, resolver, recursive=True):
        return ControlledGate(
            self._sub_gate._resolve_parameters_(resolver, recursive), self._num_controls, self._control_values
        )

    def __repr__(self):
        return f"cirq.ControlledGate({self._sub_gate!r}, num_controls={self._num_controls})"


# ---- measurement ----------------------------------------------------------------------------
class MeasurementGate(Gate):
    """Computational-basis measurement (cirq-core/cirq/ops/measurement_gate.py:30-303)."""

    def __init__(self, num_qubits: int, key: str = "", invert_mask: Sequence[bool] = ()):
        self._n = int(num_qubits)
        self._key = str(key)
        self._invert_mask = tuple(bool(b) for b in invert_mask)
        if len(self._invert_mask) > self._n:
            raise ValueError("len(invert_mask) > num_qubits")

    key = property(lambda self: self._key)
    invert_mask = property(lambda self: self._invert_mask)

    def _num_qubits_(self):
        return self._n

    def _has_unitary_(self):
        return False

    def _is_parameterized_(self):
        return False

    def full_invert_mask(self) -> tuple:
        return self._invert_mask + (False,) * (self._n - len(self._invert_mask))

    def __repr__(self):
        return f"cirq.MeasurementGate({self._n}, {self._key!r}, {self._invert_mask!r})"


def _default_measurement_key(qubits) -> str:
    return ",".join(str(q) for q in qubits)


def measure(*target, key=None, invert_mask=()) -> "GateOperation":
    """cirq.measure: the default key joins the qubits' str forms with commas."""
    if len(target) == 1 and isinstance(target[0], (list, tuple)):
        target = tuple(target[0])
    if key is None:
        key = _default_measurement_key(target)
    return MeasurementGate(len(target), key, invert_mask).on(*target)


def measure_each(*qubits, key_func=str) -> list:
    if len(qubits) == 1 and isinstance(qubits[0], (list, tuple)):
        qubits = tuple(qubits[0])
    return [MeasurementGate(1, key_func(q)).on(q) for q in qubits]


# ---- operations -----------------------------------------------------------------------------
class Operation:
    """An effect on specific qubits."""

    @property
    def qubits(self) -> tuple:
        raise NotImplementedError

    @property
    def gate(self):
        return None

    def with_qubits(self, *new_qubits) -> "Operation":
        raise NotImplementedError

    def controlled_by(self, *control_qubits, control_values=None) -> "Operation":
        g = self.gate
        if g is None:
            raise TypeError("only gate operations can be controlled")
        if not control_qubits:
            return self
        return ControlledGate(g, len(control_qubits), control_values).on(*control_qubits, *self.qubits)


class GateOperation(Operation):
    """A gate applied to qubits (cirq-core/cirq/ops/gate_operation.py)."""

    def __init__(self, gate: Gate, qubits: Sequence):
        qubits = tuple(qubits)
        n = gate.num_qubits()
        if len(qubits) != n:
            raise ValueError(f"Wrong number of qubits for <{gate!r}>. Expected {n} qubits but got <{qubits!r}>.")
        if len(set(qubits)) != len(qubits):
            raise ValueError(f"Duplicate qubits for <{gate!r}>: {qubits!r}")
        self._gate = gate
        self._qubits = qubits

    @property
    def gate(self):
        return self._gate

    @property
    def qubits(self):
        return self._qubits

    def with_qubits(self, *new_qubits):
        return GateOperation(self._gate, new_qubits)

    def with_gate(self, gate):
        return GateOperation(gate, self._qubits)

    def _unitary_(self):
        return self._gate._unitary_()

    def _has_unitary_(self):
        f = getattr(self._gate, "_has_unitary_", None)
        return bool(f()) if f else False

    def _is_parameterized_(self):
        return is_parameterized(self._gate)

    def _parameter_names_(self):
        f = getattr(self._gate, "_parameter_names_", None)
        return f() if f else set()

    def _resolve_parameters_(self, resolver, recursive=True):
        return GateOperation(self._gate._resolve_parameters_(resolver, recursive), self._qubits)

    def __pow__(self, power):
        g = self._gate.__pow__(power)
        if g is NotImplemented:
            return NotImplemented
        return GateOperation(g, self._qubits)

    def __eq__(self, other):
        if not isinstance(other, GateOperation):
            return NotImplemented
        return self._gate == other._gate and self._qubits == other._qubits

    def __hash__(self):
        return hash((self._gate, self._qubits))

    def __repr__(self):
        return f"{self._gate!r}.on({', '.join(repr(q) for q in self._qubits)})"


# ---- common instances ------------------------------------------------------------------------
H = HPowGate()
S = ZPowGate(exponent=0.5)
T = ZPowGate(exponent=0.25)
CZ = CZPowGate()
CNOT = CX = CXPowGate()
SWAP = SwapPowGate()
ISWAP = ISwapPowGate()
CCZ = CCZPowGate()
CCX = TOFFOLI = CCXPowGate()
CSWAP = FREDKIN = CSwapGate()
XX = XXPowGate()
YY = YYPowGate()
ZZ = ZZPowGate()
I = IdentityGate(1)  # noqa: E741


def identity_each(*qubits):
    return IdentityGate(len(qubits)).on(*qubits)
