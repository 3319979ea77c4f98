"""Quantum channels (Kraus / mixture) and noise models with Cirq's definitions.

Mirrors cirq-core/cirq/ops/common_channels.py and cirq-core/cirq/devices/noise_model.py.
A mixture [(p_k, U_k)] acts as the Kraus set sqrt(p_k) U_k
(cirq-core/cirq/protocols/kraus_protocol.py:194).
"""

from __future__ import annotations

import itertools
from typing import Iterable, Sequence

import numpy as np

from .gates import Gate

_I2 = np.eye(2, dtype=np.complex128)
_PX = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_PY = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
_PZ = np.array([[1, 0], [0, -1]], dtype=np.complex128)
_PMAT = {"I": _I2, "X": _PX, "Y": _PY, "Z": _PZ}


def _validate_probability(p: float, name: str) -> float:
    if p < 0:
        raise ValueError(f"{name} was less than 0.")
    if p > 1:
        raise ValueError(f"{name} was greater than 1.")
    return float(p)


class Channel(Gate):
    """A gate with no unitary: Kraus operators (and a mixture when it has one)."""

    def _has_unitary_(self):
        return False

    def _is_parameterized_(self):
        return False

    def _has_mixture_(self):
        return hasattr(self, "_mixture_")

    def _has_kraus_(self):
        return True

    def _kraus_(self) -> tuple:
        return tuple(np.sqrt(p) * np.asarray(u, dtype=np.complex128) for p, u in self._mixture_())

    def _resolve_parameters_(self, resolver, recursive=True):
        return self


class AsymmetricDepolarizingChannel(Channel):
    """Pauli-string errors with given probabilities (cirq-core/cirq/ops/common_channels.py:114-128)."""

    def __init__(self, p_x=None, p_y=None, p_z=None, error_probabilities=None, tol: float = 1e-8):
        if error_probabilities:
            n = len(next(iter(error_probabilities)))
            for k in error_probabilities:
                if len(k) != n or any(c not in "IXYZ" for c in k):
                    raise ValueError(f"bad Pauli string {k!r}")
            ident = "I" * n
            if ident not in error_probabilities:
                error_probabilities = dict(error_probabilities)
                error_probabilities[ident] = 1.0 - sum(error_probabilities.values())
            if abs(sum(error_probabilities.values()) - 1) > tol:
                raise ValueError("Probabilities do not add up to 1")
            self._num_qubits = n
            self._error_probabilities = dict(error_probabilities)
        else:
            p_x = 0.0 if p_x is None else _validate_probability(p_x, "p_x")
            p_y = 0.0 if p_y is None else _validate_probability(p_y, "p_y")
            p_z = 0.0 if p_z is None else _validate_probability(p_z, "p_z")
            p_i = 1 - _validate_probability(p_x + p_y + p_z, "p_x + p_y + p_z")
            self._num_qubits = 1
            self._error_probabilities = {"I": p_i, "X": p_x, "Y": p_y, "Z": p_z}

    @property
    def error_probabilities(self):
        return self._error_probabilities

    def _num_qubits_(self):
        return self._num_qubits

    def _mixture_(self):
        out = []
        for pauli, p in self._error_probabilities.items():
            m = np.identity(1, dtype=np.complex128)
            for c in pauli:
                m = np.kron(m, _PMAT[c])
            out.append((p, m))
        return tuple(out)

    def __repr__(self):
        return f"cirq.asymmetric_depolarize(error_probabilities={self._error_probabilities!r})"


def asymmetric_depolarize(p_x=None, p_y=None, p_z=None, error_probabilities=None, tol=1e-8):
    return AsymmetricDepolarizingChannel(p_x, p_y, p_z, error_probabilities, tol)


class DepolarizingChannel(Channel):
    """p_I = 1-p, each non-identity Pauli string p/(4^n-1) (cirq-core/cirq/ops/common_channels.py:246-300)."""

    def __init__(self, p: float, n_qubits: int = 1):
        self._p = _validate_probability(p, "p")
        self._n_qubits = int(n_qubits)
        probs = {}
        p_depol = self._p / (4**self._n_qubits - 1)
        for tup in itertools.product("IXYZ", repeat=self._n_qubits):
            s = "".join(tup)
            probs[s] = 1.0 - self._p if s == "I" * self._n_qubits else p_depol
        self._delegate = AsymmetricDepolarizingChannel(error_probabilities=probs)

    @property
    def p(self):
        return self._p

    @property
    def n_qubits(self):
        return self._n_qubits

    def _num_qubits_(self):
        return self._n_qubits

    def _mixture_(self):
        return self._delegate._mixture_()

    def __repr__(self):
        return f"cirq.depolarize(p={self._p!r})" if self._n_qubits == 1 else (
            f"cirq.depolarize(p={self._p!r}, n_qubits={self._n_qubits})"
        )


def depolarize(p: float, n_qubits: int = 1) -> DepolarizingChannel:
    return DepolarizingChannel(p, n_qubits)


class BitFlipChannel(Channel):
    """[(1-p, I), (p, X)] (cirq-core/cirq/ops/common_channels.py:1097-1098)."""

    def __init__(self, p: float):
        self._p = _validate_probability(p, "p")

    p = property(lambda self: self._p)

    def _num_qubits_(self):
        return 1

    def _mixture_(self):
        return ((1 - self._p, _I2), (self._p, _PX))

    def __repr__(self):
        return f"cirq.bit_flip(p={self._p!r})"


def bit_flip(p: float) -> BitFlipChannel:
    return BitFlipChannel(p)


class PhaseFlipChannel(Channel):
    """[(1-p, I), (p, Z)] (cirq-core/cirq/ops/common_channels.py:944-947)."""

    def __init__(self, p: float):
        self._p = _validate_probability(p, "p")

    p = property(lambda self: self._p)

    def _num_qubits_(self):
        return 1

    def _mixture_(self):
        return ((1 - self._p, _I2), (self._p, _PZ))

    def __repr__(self):
        return f"cirq.phase_flip(p={self._p!r})"


def phase_flip(p: float) -> PhaseFlipChannel:
    return PhaseFlipChannel(p)


class GeneralizedAmplitudeDampingChannel(Channel):
    """cirq-core/cirq/ops/common_channels.py:442-451."""

    def __init__(self, p: float, gamma: float):
        self._p = _validate_probability(p, "p")
        self._gamma = _validate_probability(gamma, "gamma")

    p = property(lambda self: self._p)
    gamma = property(lambda self: self._gamma)

    def _num_qubits_(self):
        return 1

    def _has_mixture_(self):
        return False

    def _kraus_(self):
        p0 = np.sqrt(self._p)
        p1 = np.sqrt(1.0 - self._p)
        sg = np.sqrt(self._gamma)
        sg1 = np.sqrt(1.0 - self._gamma)
        return (
            p0 * np.array([[1.0, 0.0], [0.0, sg1]], dtype=np.complex128),
            p0 * np.array([[0.0, sg], [0.0, 0.0]], dtype=np.complex128),
            p1 * np.array([[sg1, 0.0], [0.0, 1.0]], dtype=np.complex128),
            p1 * np.array([[0.0, 0.0], [sg, 0.0]], dtype=np.complex128),
        )

    def __repr__(self):
        return f"cirq.generalized_amplitude_damp(p={self._p!r},gamma={self._gamma!r})"


def generalized_amplitude_damp(p: float, gamma: float) -> GeneralizedAmplitudeDampingChannel:
    return GeneralizedAmplitudeDampingChannel(p, gamma)


class AmplitudeDampingChannel(Channel):
    """The first two Kraus operators of GAD(1, gamma) (cirq-core/cirq/ops/common_channels.py:578-581)."""

    def __init__(self, gamma: float):
        self._gamma = _validate_probability(gamma, "gamma")
        self._delegate = GeneralizedAmplitudeDampingChannel(1.0, self._gamma)

    gamma = property(lambda self: self._gamma)

    def _num_qubits_(self):
        return 1

    def _has_mixture_(self):
        return False

    def _kraus_(self):
        return tuple(self._delegate._kraus_())[:2]

    def __repr__(self):
        return f"cirq.amplitude_damp(gamma={self._gamma!r})"


def amplitude_damp(gamma: float) -> AmplitudeDampingChannel:
    return AmplitudeDampingChannel(gamma)


class PhaseDampingChannel(Channel):
    """[[1,0],[0,sqrt(1-g)]], [[0,0],[0,sqrt(g)]] (cirq-core/cirq/ops/common_channels.py:836-840)."""

    def __init__(self, gamma: float):
        self._gamma = _validate_probability(gamma, "gamma")

    gamma = property(lambda self: self._gamma)

    def _num_qubits_(self):
        return 1

    def _has_mixture_(self):
        return False

    def _kraus_(self):
        return (
            np.array([[1.0, 0.0], [0.0, np.sqrt(1.0 - self._gamma)]], dtype=np.complex128),
            np.array([[0.0, 0.0], [0.0, np.sqrt(self._gamma)]], dtype=np.complex128),
        )

    def __repr__(self):
        return f"cirq.phase_damp(gamma={self._gamma!r})"


def phase_damp(gamma: float) -> PhaseDampingChannel:
    return PhaseDampingChannel(gamma)


class ResetChannel(Channel):
    """Reset to |0>: Kraus |0><0|, |0><1| (cirq-core/cirq/ops/common_channels.py:712-716)."""

    def _num_qubits_(self):
        return 1

    def _has_mixture_(self):
        return False

    def _kraus_(self):
        k0 = np.zeros((2, 2), dtype=np.complex128)
        k1 = np.zeros((2, 2), dtype=np.complex128)
        k0[0, 0] = 1
        k1[0, 1] = 1
        return (k0, k1)

    def __repr__(self):
        return "cirq.ResetChannel()"


def reset(qubit):
    return ResetChannel().on(qubit)


class KrausChannel(Channel):
    """A channel given by explicit Kraus operators (cirq.KrausChannel)."""

    def __init__(self, kraus_ops: Sequence, validate: bool = False):
        ops = tuple(np.asarray(k, dtype=np.complex128) for k in kraus_ops)
        if not ops:
            raise ValueError("KrausChannel needs at least one operator")
        d = ops[0].shape[0]
        if any(k.shape != (d, d) for k in ops) or d & (d - 1):
            raise ValueError("Kraus operators must be 2^k x 2^k and of one size")
        if validate:
            s = sum(k.conj().T @ k for k in ops)
            if not np.allclose(s, np.eye(d), atol=1e-6):
                raise ValueError("Kraus operators do not form a CPTP map")
        self._ops = ops

    def _num_qubits_(self):
        return int(round(np.log2(self._ops[0].shape[0])))

    def _has_mixture_(self):
        return False

    def _kraus_(self):
        return self._ops


class MixedUnitaryChannel(Channel):
    """A channel given by (probability, unitary) pairs (cirq.MixedUnitaryChannel)."""

    def __init__(self, mixture: Iterable):
        self._mix = tuple((float(p), np.asarray(u, dtype=np.complex128)) for p, u in mixture)

    def _num_qubits_(self):
        return int(round(np.log2(self._mix[0][1].shape[0])))

    def _mixture_(self):
        return self._mix


# ---- noise models -------------------------------------------------------------------------------
class NoiseModel:
    """Base noise model: turns each moment into noisy moments (cirq-core/cirq/devices/noise_model.py:30-120)."""

    def noisy_moments(self, moments, system_qubits) -> list:
        out = []
        for m in moments:
            r = self.noisy_moment(m, system_qubits)
            out.extend(r if isinstance(r, (list, tuple)) else [r])
        return out

    def noisy_moment(self, moment, system_qubits):
        return moment

    @staticmethod
    def from_noise_model_like(noise) -> "NoiseModel":
        if noise is None:
            return NO_NOISE
        if isinstance(noise, NoiseModel):
            return noise
        if isinstance(noise, Gate):
            if noise.num_qubits() != 1:
                raise ValueError("Multi-qubit gates cannot be used as noise")
            return ConstantQubitNoiseModel(noise)
        raise TypeError(f"Expected a NOISE_MODEL_LIKE, got {noise!r}")


class _NoNoise(NoiseModel):
    pass


NO_NOISE = _NoNoise()


class ConstantQubitNoiseModel(NoiseModel):
    """Appends a moment of the noise gate on every system qubit (cirq-core/cirq/devices/noise_model.py:239-253)."""

    def __init__(self, qubit_noise_gate: Gate, prepend: bool = False):
        if qubit_noise_gate.num_qubits() != 1:
            raise ValueError("noise.num_qubits() != 1")
        self.qubit_noise_gate = qubit_noise_gate
        self._prepend = prepend

    def noisy_moment(self, moment, system_qubits):
        from .circuits import Moment

        if getattr(moment, "_is_noise", False):
            return moment
        noise = Moment([self.qubit_noise_gate.on(q) for q in system_qubits])
        noise._is_noise = True
        out = [moment, noise]
        return out[::-1] if self._prepend else out
