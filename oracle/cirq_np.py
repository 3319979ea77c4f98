"""NumPy restatement of Cirq's state-vector / density-matrix simulation (test oracle).

Ops are neutral tuples ``(kind, family, params, axes)``:
  ('gate', family, params, axes)     a unitary gate on qubit axes (Cirq big-endian axes)
  ('channel', family, params, axes)  a quantum channel (density-matrix runs only)
Families and their matrices restate the reference's definitions (file:line on each table).

Functions and the reference code they follow:
  apply_matrix_to_slices      cirq-core/cirq/linalg/transformations.py:310-380
  targeted_left_multiply      cirq-core/cirq/linalg/transformations.py:105-172
  apply_unitary               cirq-core/cirq/protocols/apply_unitary_protocol.py:440-465
  apply_channel               cirq-core/cirq/protocols/apply_channel_protocol.py:274-356
  simulate_state_vector       cirq-core/cirq/sim/simulator_base.py:169-214 (op loop, no splitting)
  simulate_density_matrix     cirq-core/cirq/sim/density_matrix_simulation_state.py:161-188
  sample_state_vector         cirq-core/cirq/sim/state_vector.py:170-232
  state_probabilities_by_indices  cirq-core/cirq/sim/simulation_utils.py:24-65
  sample_density_matrix       cirq-core/cirq/sim/density_matrix_utils.py:31-94,185-192
  expectation_from_state_vector   cirq-core/cirq/ops/pauli_string.py:625-656
  expectation_from_density_matrix cirq-core/cirq/ops/pauli_string.py:734-764
"""

from __future__ import annotations

import itertools
import math

import numpy as np

# ------------------------------------------------------------------------------ gate tables
# EigenGate: U = sum_k P_k * 1j**(2 t (h_k + s))   (cirq-core/cirq/ops/eigen_gate.py:295-305)
_R2 = math.sqrt(2)
EIGEN = {
    # cirq-core/cirq/ops/common_gates.py:146-156 (dimension 2)
    "XPowGate": [(0, [[0.5, 0.5], [0.5, 0.5]]), (1, [[0.5, -0.5], [-0.5, 0.5]])],
    # cirq-core/cirq/ops/common_gates.py:474-478
    "YPowGate": [(0, [[0.5, -0.5j], [0.5j, 0.5]]), (1, [[0.5, 0.5j], [-0.5j, 0.5]])],
    # cirq-core/cirq/ops/common_gates.py:750-759
    "ZPowGate": [(0, [[1, 0], [0, 0]]), (1, [[0, 0], [0, 1]])],
    # cirq-core/cirq/ops/common_gates.py:947-955
    "HPowGate": [
        (0, (np.array([[3 + 2 * _R2, 1 + _R2], [1 + _R2, 1]]) / (4 + 2 * _R2)).tolist()),
        (1, (np.array([[3 - 2 * _R2, 1 - _R2], [1 - _R2, 1]]) / (4 - 2 * _R2)).tolist()),
    ],
    # cirq-core/cirq/ops/common_gates.py:1064-1065
    "CZPowGate": [(0, np.diag([1, 1, 1, 0]).tolist()), (1, np.diag([0, 0, 0, 1]).tolist())],
    # cirq-core/cirq/ops/common_gates.py (CXPowGate._eigen_components)
    "CXPowGate": [
        (0, [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0.5, 0.5], [0, 0, 0.5, 0.5]]),
        (1, [[0, 0, 0, 0], [0, 0, 0, 0], [0, 0, 0.5, -0.5], [0, 0, -0.5, 0.5]]),
    ],
    # cirq-core/cirq/ops/swap_gates.py:86-97
    "SwapPowGate": [
        (0, [[1, 0, 0, 0], [0, 0.5, 0.5, 0], [0, 0.5, 0.5, 0], [0, 0, 0, 1]]),
        (1, [[0, 0, 0, 0], [0, 0.5, -0.5, 0], [0, -0.5, 0.5, 0], [0, 0, 0, 0]]),
    ],
    # cirq-core/cirq/ops/swap_gates.py:207-219
    "ISwapPowGate": [
        (0, np.diag([1, 0, 0, 1]).tolist()),
        (0.5, [[0, 0, 0, 0], [0, 0.5, 0.5, 0], [0, 0.5, 0.5, 0], [0, 0, 0, 0]]),
        (-0.5, [[0, 0, 0, 0], [0, 0.5, -0.5, 0], [0, -0.5, 0.5, 0], [0, 0, 0, 0]]),
    ],
    # cirq-core/cirq/ops/parity_gates.py:69-81
    "XXPowGate": [
        (0, [[0.5, 0, 0, 0.5], [0, 0.5, 0.5, 0], [0, 0.5, 0.5, 0], [0.5, 0, 0, 0.5]]),
        (1, [[0.5, 0, 0, -0.5], [0, 0.5, -0.5, 0], [0, -0.5, 0.5, 0], [-0.5, 0, 0, 0.5]]),
    ],
    # cirq-core/cirq/ops/parity_gates.py:155-169
    "YYPowGate": [
        (0, [[0.5, 0, 0, -0.5], [0, 0.5, 0.5, 0], [0, 0.5, 0.5, 0], [-0.5, 0, 0, 0.5]]),
        (1, [[0.5, 0, 0, 0.5], [0, 0.5, -0.5, 0], [0, -0.5, 0.5, 0], [0.5, 0, 0, 0.5]]),
    ],
    # cirq-core/cirq/ops/parity_gates.py:237-238
    "ZZPowGate": [(0, np.diag([1, 0, 0, 1]).tolist()), (1, np.diag([0, 1, 1, 0]).tolist())],
    # cirq-core/cirq/ops/three_qubit_gates.py:63-64
    "CCZPowGate": [(0, np.diag([1, 1, 1, 1, 1, 1, 1, 0]).tolist()), (1, np.diag([0] * 7 + [1]).tolist())],
}


def _ccx_components():
    # cirq-core/cirq/ops/three_qubit_gates.py:394-403
    c0 = np.zeros((8, 8))
    c0[:6, :6] = np.eye(6)
    c0[6:, 6:] = [[0.5, 0.5], [0.5, 0.5]]
    c1 = np.zeros((8, 8))
    c1[6:, 6:] = [[0.5, -0.5], [-0.5, 0.5]]
    return [(0, c0), (1, c1)]


def eigen_unitary(family: str, exponent: float, global_shift: float = 0.0) -> np.ndarray:
    comps = _ccx_components() if family == "CCXPowGate" else EIGEN[family]
    return np.sum(
        [np.asarray(c, dtype=np.complex128) * 1j ** (2 * exponent * (h + global_shift)) for h, c in comps], axis=0
    )


def gate_unitary(family: str, params: dict) -> np.ndarray:
    """Matrix of a gate family with Cirq's exact global phase."""
    if family in EIGEN or family == "CCXPowGate":
        return eigen_unitary(family, params.get("exponent", 1.0), params.get("global_shift", 0.0))
    if family in ("Rx", "Ry", "Rz"):
        base = {"Rx": "XPowGate", "Ry": "YPowGate", "Rz": "ZPowGate"}[family]
        return eigen_unitary(base, params["rads"] / np.pi, -0.5)
    if family == "PhasedXPowGate":  # cirq-core/cirq/ops/phased_x_gate.py:138-145
        z = eigen_unitary("ZPowGate", params["phase_exponent"])
        x = eigen_unitary("XPowGate", params["exponent"])
        p = np.exp(1j * np.pi * params.get("global_shift", 0.0) * params["exponent"])
        return np.dot(np.dot(z, x), np.conj(z)) * p
    if family == "PhasedXZGate":  # decomposition of cirq-core/cirq/ops/phased_x_z_gate.py:209-213
        a, x, z = params["axis_phase_exponent"], params["x_exponent"], params["z_exponent"]
        return (
            eigen_unitary("ZPowGate", a + z) @ eigen_unitary("XPowGate", x) @ eigen_unitary("ZPowGate", -a)
        )
    if family == "FSimGate":  # cirq-core/cirq/ops/fsim_gate.py:135-150
        a = math.cos(params["theta"])
        b = -1j * math.sin(params["theta"])
        c = np.exp(-1j * params["phi"])
        return np.array([[1, 0, 0, 0], [0, a, b, 0], [0, b, a, 0], [0, 0, 0, c]], dtype=np.complex128)
    if family == "DiagonalGate":  # cirq-core/cirq/ops/diagonal_gate.py:100-122
        return np.diag(np.exp(1j * np.asarray(params["angles"], dtype=np.float64)))
    if family in ("MatrixGate", "Unitary"):
        return np.asarray(params["matrix"], dtype=np.complex128)
    if family == "CSwapGate":
        u = np.eye(8, dtype=np.complex128)
        u[[5, 6]] = u[[6, 5]]
        return u
    if family == "IdentityGate":
        return np.eye(2 ** params.get("num_qubits", 1), dtype=np.complex128)
    if family == "Pauli":
        return {"X": np.array([[0, 1], [1, 0]]), "Y": np.array([[0, -1j], [1j, 0]]), "Z": np.diag([1, -1])}[
            params["pauli"]
        ].astype(np.complex128)
    if family == "Controlled":
        sub = gate_unitary(params["sub_family"], params["sub_params"])
        nc = len(params["control_values"])
        ds = sub.shape[0]
        u = np.eye(ds * 2**nc, dtype=np.complex128)
        ctl = 0
        for v in params["control_values"]:
            ctl = ctl * 2 + int(v)
        u[ctl * ds : (ctl + 1) * ds, ctl * ds : (ctl + 1) * ds] = sub
        return u
    raise KeyError(f"oracle has no definition for gate family {family!r}")


_PX = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_PY = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
_PZ = np.array([[1, 0], [0, -1]], dtype=np.complex128)
_PI = np.eye(2, dtype=np.complex128)
_PM = {"I": _PI, "X": _PX, "Y": _PY, "Z": _PZ}


def channel_kraus(family: str, params: dict) -> list:
    """Kraus operators (mixtures become sqrt(p) U, cirq-core/cirq/protocols/kraus_protocol.py:194)."""
    if family == "depolarize":  # cirq-core/cirq/ops/common_channels.py:281-293
        n = params.get("n_qubits", 1)
        p = params["p"]
        out = []
        for tup in itertools.product("IXYZ", repeat=n):
            m = np.identity(1, dtype=np.complex128)
            for c in tup:
                m = np.kron(m, _PM[c])
            pk = 1 - p if set(tup) == {"I"} else p / (4**n - 1)
            out.append(np.sqrt(pk) * m)
        return out
    if family == "asymmetric_depolarize":  # cirq-core/cirq/ops/common_channels.py:114-128
        return [np.sqrt(p) * u for p, u in channel_mixture(family, params)]
    if family == "bit_flip":
        p = params["p"]
        return [np.sqrt(1 - p) * _PI, np.sqrt(p) * _PX]
    if family == "phase_flip":
        p = params["p"]
        return [np.sqrt(1 - p) * _PI, np.sqrt(p) * _PZ]
    if family in ("amplitude_damp", "generalized_amplitude_damp"):  # cirq-core/cirq/ops/common_channels.py:442-451
        g = params["gamma"]
        p = params.get("p", 1.0)
        ops = [
            np.sqrt(p) * np.array([[1, 0], [0, np.sqrt(1 - g)]]),
            np.sqrt(p) * np.array([[0, np.sqrt(g)], [0, 0]]),
            np.sqrt(1 - p) * np.array([[np.sqrt(1 - g), 0], [0, 1]]),
            np.sqrt(1 - p) * np.array([[0, 0], [np.sqrt(g), 0]]),
        ]
        ops = [o.astype(np.complex128) for o in ops]
        return ops[:2] if family == "amplitude_damp" else ops  # cirq-core/cirq/ops/common_channels.py:578-581
    if family == "phase_damp":  # cirq-core/cirq/ops/common_channels.py:836-840
        g = params["gamma"]
        return [np.array([[1, 0], [0, np.sqrt(1 - g)]], dtype=np.complex128),
                np.array([[0, 0], [0, np.sqrt(g)]], dtype=np.complex128)]
    if family == "reset":
        return [np.array([[1, 0], [0, 0]], dtype=np.complex128), np.array([[0, 1], [0, 0]], dtype=np.complex128)]
    if family == "kraus":
        return [np.asarray(k, dtype=np.complex128) for k in params["ops"]]
    raise KeyError(f"oracle has no definition for channel family {family!r}")


# ------------------------------------------------------------------------------ linalg kernels
def slice_for_qubits_equal_to(axes, value: int, n: int):
    """Index selecting axes == bits of value (big-endian over axes); cirq.slice_for_qubits_equal_to."""
    idx = [slice(None)] * n
    k = len(axes)
    for j, a in enumerate(axes):
        idx[a] = (value >> (k - 1 - j)) & 1
    return tuple(idx)


def apply_matrix_to_slices(target, matrix, slices, out=None):
    if out is target:
        raise ValueError("Can't write output over the input.")
    if out is None:
        out = np.copy(target)
    else:
        out[...] = target[...]
    for i, s_i in enumerate(slices):
        out[s_i] *= matrix[i, i]
        for j, s_j in enumerate(slices):
            if i != j:
                out[s_i] += target[s_j] * matrix[i, j]
    return out


def targeted_left_multiply(left_matrix, right_target, target_axes, out=None):
    k = len(target_axes)
    d = len(right_target.shape)
    work = tuple(range(k))
    data = tuple(range(k, k + d))
    used = tuple(data[q] for q in target_axes)
    inp = work + used
    outi = list(data)
    for w, t in zip(work, target_axes):
        outi[t] = w
    all_idx = set(inp + data + tuple(outi))
    return np.einsum(left_matrix, inp, right_target, data, outi, optimize=len(all_idx) >= 26, out=out)


def apply_unitary(tensor: np.ndarray, matrix: np.ndarray, axes) -> np.ndarray:
    """New tensor = matrix applied to `axes` (1q: slices; otherwise einsum)."""
    matrix = matrix.astype(tensor.dtype)
    k = len(axes)
    if k == 1:
        idx0 = [slice(None)] * tensor.ndim
        idx1 = [slice(None)] * tensor.ndim
        idx0[axes[0]] = 0
        idx1[axes[0]] = 1
        return apply_matrix_to_slices(tensor, matrix, [tuple(idx0), tuple(idx1)])
    return targeted_left_multiply(matrix.reshape((2,) * (2 * k)), tensor, list(axes))


def apply_channel(rho: np.ndarray, kraus: list, axes, n: int) -> np.ndarray:
    """sum_k K rho K^dagger on left axes `axes` and right axes `axes + n`."""
    left = list(axes)
    right = [a + n for a in axes]
    if len(kraus) == 1:
        r = apply_unitary(rho, kraus[0], left)
        r = np.conjugate(r)
        r = apply_unitary(r, kraus[0], right)
        return np.conjugate(r)
    out = np.zeros_like(rho)
    for K in kraus:
        K = K.astype(rho.dtype)
        t = apply_unitary(rho, K, left)
        t = apply_unitary(t, np.conjugate(K), right)
        out += t
    return out


# ------------------------------------------------------------------------------ gates' own strategies
def subspace_index(axes, little_endian_bits: int, ndim: int):
    """Index selecting the target axes equal to the bits of `little_endian_bits`, the LEAST
    significant bit for the FIRST axis (cirq-core/cirq/protocols/apply_unitary_protocol.py:184-214)."""
    idx = [slice(None)] * ndim
    for j, a in enumerate(axes):
        idx[a] = (little_endian_bits >> j) & 1
    return tuple(idx)


def _global_phase(exponent: float, global_shift: float):
    return 1j ** (2 * exponent * global_shift)


def gate_apply_unitary(t: np.ndarray, family: str, params: dict, axes):
    """A gate's own ``_apply_unitary_`` where Cirq defines one for the exponents it accepts; None
    (NotImplemented) otherwise, and the caller falls back to the matrix strategy.

    XPowGate   cirq-core/cirq/ops/common_gates.py:123-133 (exponent 1: swap of the slices)
    YPowGate   cirq-core/cirq/ops/common_gates.py:454-464 (exponent 1)
    ZPowGate   cirq-core/cirq/ops/common_gates.py:658-669 (any exponent: phase on |1>); Rz is one
    HPowGate   cirq-core/cirq/ops/common_gates.py:977-987 (exponent 1)
    CZPowGate  cirq-core/cirq/ops/common_gates.py:1072-1082 (any exponent: phase on |11>)
    CXPowGate  cirq-core/cirq/ops/common_gates.py:1257-1268 (exponent 1)
    ZZPowGate  cirq-core/cirq/ops/parity_gates.py:253-265 (any exponent: phase on |01>, |10>)
    The result is a new array when Cirq writes its ``available_buffer``, else ``t`` in place.
    """
    nd = t.ndim
    axes = list(axes)
    if family == "Rz":
        family, params = "ZPowGate", {"exponent": params["rads"] / np.pi, "global_shift": -0.5}
    e = float(params.get("exponent", 1.0))
    s = float(params.get("global_shift", 0.0))
    p = _global_phase(e, s)
    if family in ("XPowGate", "YPowGate") and e == 1:
        zero, one = subspace_index(axes, 0, nd), subspace_index(axes, 1, nd)
        buf = np.empty_like(t)
        if family == "XPowGate":
            buf[zero] = t[one]
            buf[one] = t[zero]
        else:
            buf[zero] = -1j * t[one]
            buf[one] = 1j * t[zero]
        if p != 1:
            buf *= p
        return buf
    if family == "ZPowGate":
        t[subspace_index(axes, 1, nd)] *= 1j ** (e * 4 * 1 / 2)
        if p != 1:
            t *= p
        return t
    if family == "HPowGate" and e == 1:
        zero, one = subspace_index(axes, 0, nd), subspace_index(axes, 1, nd)
        t[one] -= t[zero]
        t[one] *= -0.5
        t[zero] -= t[one]
        t *= np.sqrt(2) * p
        return t
    if family == "CZPowGate":
        t[subspace_index(axes, 0b11, nd)] *= 1j ** (2 * e)
        if p != 1:
            t *= p
        return t
    if family == "CXPowGate" and e == 1:
        oo, zo = subspace_index(axes, 0b11, nd), subspace_index(axes, 0b01, nd)
        keep = np.copy(t[oo])
        t[oo] = t[zo]
        t[zo] = keep
        if p != 1:
            t *= p
        return t
    if family == "ZZPowGate":
        if p != 1:
            t *= p
        rel = 1j ** (2 * e)
        t[subspace_index(axes, 0b10, nd)] *= rel
        t[subspace_index(axes, 0b01, nd)] *= rel
        return t
    return None


def apply_gate(t: np.ndarray, family: str, params: dict, axes) -> np.ndarray:
    """Cirq's strategy order for a gate (cirq-core/cirq/protocols/apply_unitary_protocol.py:375-389):
    the gate's own ``_apply_unitary_`` first, then its matrix."""
    r = gate_apply_unitary(t, family, params, axes)
    if r is not None:
        return r
    return apply_unitary(t, gate_unitary(family, params), list(axes))


# ------------------------------------------------------------------------------ simulation
def op_matrix(op) -> np.ndarray:
    kind, family, params, axes = op
    return gate_unitary(family, params)


def simulate_state_vector(ops, n: int, dtype=np.complex128, initial_state: int = 0) -> np.ndarray:
    psi = np.zeros(2**n, dtype=dtype)
    psi[initial_state] = 1
    t = psi.reshape((2,) * n)
    for op in ops:
        kind, family, params, axes = op
        if kind != "gate":
            raise ValueError("state-vector oracle runs take gates only")
        if family == "GlobalPhase":
            t = t * params["coefficient"]
            continue
        t = apply_gate(t, family, params, axes)
    return t.reshape(-1)


def simulate_density_matrix(ops, n: int, dtype=np.complex128, initial_state: int = 0) -> np.ndarray:
    rho = np.zeros((2**n, 2**n), dtype=dtype)
    rho[initial_state, initial_state] = 1
    t = rho.reshape((2,) * (2 * n))
    for op in ops:
        kind, family, params, axes = op
        if kind == "gate":
            if family == "GlobalPhase":
                continue
            t = apply_channel(t, [gate_unitary(family, params)], list(axes), n)
        else:
            t = apply_channel(t, channel_kraus(family, params), list(axes), n)
    return t.reshape(2**n, 2**n)


# ------------------------------------------------------------------------------ sampling
def big_endian_int_to_digits(val: int, base) -> list:
    """cirq-core/cirq/value/digits.py:125-170 (digit_count = len(base))."""
    out = []
    for b in reversed(base):
        val, r = divmod(val, b)
        out.append(r)
    if val:
        raise ValueError("Out of range.")
    return out[::-1]


def state_probabilities_by_indices(state_probability, indices, qid_shape):
    probs = state_probability.reshape((-1,))
    not_measured = [i for i in range(len(qid_shape)) if i not in indices]
    probs = probs.reshape(qid_shape)
    probs = np.transpose(probs, list(indices) + not_measured)
    probs = probs.reshape((-1,))
    if len(not_measured):
        volume = np.prod([qid_shape[i] for i in indices])
        probs = probs.reshape((volume, -1))
        probs = np.sum(probs, axis=-1)
    probs = np.clip(probs, 0, None)
    probs /= probs.sum()
    return probs


def state_vector_probs(state_vector, indices):
    n = int(round(math.log2(state_vector.size)))
    probs = (state_vector * state_vector.conj()).real
    return state_probabilities_by_indices(probs, indices, (2,) * n)


def sample_state_vector(state_vector, indices, repetitions: int, prng: np.random.RandomState) -> np.ndarray:
    if repetitions == 0 or len(indices) == 0:
        return np.zeros((repetitions, len(indices)), dtype=np.uint8)
    probs = state_vector_probs(state_vector, indices)
    result = prng.choice(len(probs), size=repetitions, p=probs)
    shape = (2,) * len(indices)
    return np.array([big_endian_int_to_digits(int(r), shape) for r in result], dtype=np.uint8)


def density_matrix_probs(rho, indices):
    n = int(round(math.log2(rho.shape[0])))
    all_probs = np.diagonal(rho.reshape(2**n, 2**n))
    return state_probabilities_by_indices(all_probs.real, indices, (2,) * n)


def sample_density_matrix(rho, indices, repetitions: int, prng: np.random.RandomState) -> np.ndarray:
    probs = density_matrix_probs(rho, indices)
    result = prng.choice(len(probs), size=repetitions, p=probs)
    shape = (2,) * len(indices)
    return np.array([big_endian_int_to_digits(int(r), shape) for r in result], dtype=np.uint8)


# ------------------------------------------------------------------------------ expectations
def expectation_from_state_vector(state, term, n: int) -> complex:
    """term = (coefficient, {axis: 'X'|'Y'|'Z'})."""
    coef, paulis = term
    t = state.reshape((2,) * n)
    ket = np.copy(t)
    for a, p in paulis.items():
        ket = apply_unitary(ket, _PM[p], [a])
    return coef * np.tensordot(t.conj(), ket, axes=n).item()


def expectation_from_density_matrix(rho, term, n: int) -> float:
    coef, paulis = term
    r = np.copy(rho).reshape((2,) * (2 * n))
    for a, p in paulis.items():
        r = apply_unitary(r, _PM[p], [a])
    while any(r.shape):
        r = np.trace(r, axis1=0, axis2=len(r.shape) // 2)
    return float(np.real(r * coef))


# ------------------------------------------------------------------------------ trajectories
def channel_mixture(family: str, params: dict):
    """cirq.mixture of the mixture channels, [(p_i, U_i)] in Cirq's order, or None for channels
    that only have Kraus operators.
      depolarize             cirq-core/cirq/ops/common_channels.py:281-293 (I, then Paulis)
      asymmetric_depolarize  cirq-core/cirq/ops/common_channels.py:114-128
      bit_flip / phase_flip  cirq-core/cirq/ops/common_channels.py (BitFlipChannel._mixture_,
                             PhaseFlipChannel._mixture_: identity first)"""
    if family == "depolarize":
        n = params.get("n_qubits", 1)
        p = params["p"]
        out = []
        for tup in itertools.product("IXYZ", repeat=n):
            m = np.identity(1, dtype=np.complex128)
            for c in tup:
                m = np.kron(m, _PM[c])
            out.append((1 - p if set(tup) == {"I"} else p / (4**n - 1), m))
        return out
    if family == "asymmetric_depolarize":
        ep = params.get("error_probabilities")
        if ep is None:
            px, py, pz = params.get("p_x", 0.0), params.get("p_y", 0.0), params.get("p_z", 0.0)
            ep = {"I": 1 - px - py - pz, "X": px, "Y": py, "Z": pz}
        out = []
        for pauli, p in ep.items():  # the channel's own order of Pauli strings
            m = np.identity(1, dtype=np.complex128)
            for c in pauli:
                m = np.kron(m, _PM[c])
            out.append((p, m))
        return out
    if family == "bit_flip":
        return [(1 - params["p"], _PI), (params["p"], _PX)]
    if family == "phase_flip":
        return [(1 - params["p"], _PI), (params["p"], _PZ)]
    return None


def split_prefix(moments, in_prefix):
    """split_into_matching_protocol_then_general (cirq-core/cirq/sim/simulator.py:952-983) on
    moments of neutral ops: an op joins the prefix when `in_prefix(op)` holds and no earlier op
    sent to the suffix shares a qubit (or, for measurements, a key) with it."""
    blocked_q, blocked_k = set(), set()
    prefix, suffix = [], []
    for moment in moments:
        mp, gp = [], []
        for op in moment:
            qs = set(op[3])
            keys = {op[1]} if op[0] == "measure" else set()
            if in_prefix(op) and qs.isdisjoint(blocked_q) and keys.isdisjoint(blocked_k):
                mp.append(op)
            else:
                blocked_q |= qs
                blocked_k |= keys
                gp.append(op)
        if mp:
            prefix.append(mp)
        if gp:
            suffix.append(gp)
    return prefix, suffix


def _measure(t: np.ndarray, axes, invert, prng):
    """SimulationState.measure -> measure_state_vector (cirq-core/cirq/sim/state_vector.py:235-322):
    probabilities in the state's precision, one prng.choice, projection, renormalisation, then
    the invert mask (cirq-core/cirq/sim/simulation_state.py measure)."""
    n = t.ndim
    probs = (t * t.conj()).real
    probs = state_probabilities_by_indices(probs, list(axes), (2,) * n)
    result = int(prng.choice(len(probs), p=probs))
    bits = big_endian_int_to_digits(result, (2,) * len(axes))
    mask = np.ones(t.shape, dtype=bool)
    mask[slice_for_qubits_equal_to(list(axes), result, n)] = False
    out = np.copy(t)
    out[mask] = 0
    out /= np.sqrt(probs[result])
    bits = [b ^ int(bool(invert[i])) if i < len(invert) else b for i, b in enumerate(bits)]
    return bits, out


def _apply_stochastic_channel(t: np.ndarray, family: str, params: dict, axes, prng) -> np.ndarray:
    """act_on for a channel on a state vector (cirq-core/cirq/sim/state_vector_simulation_state.py:
    183-257, strategy order :376-399): a mixture draws its unitary with prng.choice; otherwise
    one prng.random() walks the Kraus operators by their weights ||K_i psi||^2, falling back to the
    heaviest branch when rounding leaves the draw unassigned."""
    mix = channel_mixture(family, params)
    if mix is not None:
        probs, unitaries = zip(*mix)
        index = prng.choice(range(len(unitaries)), p=probs)
        return apply_unitary(t, np.asarray(unitaries[index]), list(axes))
    kraus = channel_kraus(family, params)
    p = prng.random()
    fallback_weight, fallback_index = 0.0, 0
    buf = None
    for index, k in enumerate(kraus):
        buf = apply_unitary(t, k, list(axes))
        weight = float(np.linalg.norm(buf) ** 2)
        if weight > fallback_weight:
            fallback_index, fallback_weight = index, weight
        p -= weight
        if p < 0:
            break
    if p >= 0 or weight == 0:
        buf = apply_unitary(t, kraus[fallback_index], list(axes))
        weight = fallback_weight
    return buf / np.sqrt(weight)


def run_trajectories(moments, n: int, repetitions: int, prng: np.random.RandomState, dtype=np.complex64) -> dict:
    """SimulatorBase._run with a general suffix (cirq-core/cirq/sim/simulator_base.py:216-273) on
    one merged state vector: the unitary prefix once, then per repetition, in order, the suffix's
    gates, measurements (records per key, one row per instance) and channels on a copy of the
    prefix state; records padded as pad_evenly does.  `moments`: lists of neutral ops, with
    ('measure', key, {'invert': mask}, axes) for measurements."""
    prefix, suffix = split_prefix(moments, lambda op: op[0] == "gate")
    psi = np.zeros(2**n, dtype=dtype)
    psi[0] = 1
    t0 = psi.reshape((2,) * n)
    for m in prefix:
        for _, fam, params, axes in m:
            t0 = apply_gate(t0, fam, params, axes) if fam != "GlobalPhase" else t0 * params["coefficient"]
    records: dict = {}
    for _ in range(repetitions):
        t = np.copy(t0)
        rec: dict = {}
        for m in suffix:
            for kind, fam, params, axes in m:
                if kind == "measure":
                    bits, t = _measure(t, axes, params.get("invert", ()), prng)
                    rec.setdefault(fam, []).append(bits)
                elif kind == "channel":
                    t = _apply_stochastic_channel(t, fam, params, axes, prng)
                elif fam == "GlobalPhase":
                    t = t * params["coefficient"]
                else:
                    t = apply_gate(t, fam, params, axes)
        for k, r in rec.items():
            records.setdefault(k, []).append(r)
    out = {}
    for k, v in records.items():
        largest = max(len(r) for r in v)
        xs = np.zeros((len(v), largest, len(v[0][0])), dtype=np.uint8)
        for i, r in enumerate(v):
            xs[i, : len(r), :] = r
        out[k] = xs
    return out
