"""Pins of the oracle's gate and channel tables against the reference tests' literal known answers.

Each test quotes the reference test it restates (file:line under cirq-core/cirq/).  The answers are
literal matrices or Kraus lists written in those tests, so a swapped projector, a wrong sign in an
eigenphase, a transposed operand or a reordered Kraus list fails here.  Kraus order matters beyond
the channel's action: trajectory runs draw a branch by walking the list in order.
"""

import itertools

import numpy as np
import pytest

from oracle import cirq_np as O

X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
Z = np.diag([1, -1]).astype(np.complex128)
I2 = np.eye(2, dtype=np.complex128)


def U(family, **params):
    return O.gate_unitary(family, params)


def kraus(family, **params):
    return [np.asarray(k) for k in O.channel_kraus(family, params)]


def assert_kraus(got, want):
    assert len(got) == len(want)
    for g, w in zip(got, want):
        np.testing.assert_allclose(g, np.asarray(w, dtype=np.complex128), atol=1e-12)


def allclose_up_to_global_phase(a, b, atol=1e-8):
    k = np.unravel_index(np.argmax(np.abs(b)), b.shape)
    ph = a[k] / b[k]
    return abs(abs(ph) - 1) < atol and np.allclose(a, ph * b, atol=atol)


# ------------------------------------------------------------------------------ parity gates
def test_xx_matrix_known_answers():
    # ops/parity_gates_test.py:71-84
    np.testing.assert_allclose(U("XXPowGate"), np.fliplr(np.eye(4)), atol=1e-12)
    np.testing.assert_allclose(U("XXPowGate", exponent=2.0), np.eye(4), atol=1e-12)
    c, s = np.cos(np.pi / 6), -1j * np.sin(np.pi / 6)
    want = np.array([[c, 0, 0, s], [0, c, s, 0], [0, s, c, 0], [s, 0, 0, c]])
    np.testing.assert_allclose(U("XXPowGate", exponent=1 / 3, global_shift=-0.5), want, atol=1e-12)


def test_yy_matrix_known_answers():
    # ops/parity_gates_test.py:143-156
    want = np.array([[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]])
    np.testing.assert_allclose(U("YYPowGate"), want, atol=1e-12)
    np.testing.assert_allclose(U("YYPowGate", exponent=2.0), np.eye(4), atol=1e-12)
    c, s = np.cos(np.pi / 6), 1j * np.sin(np.pi / 6)
    want = np.array([[c, 0, 0, s], [0, c, -s, 0], [0, -s, c, 0], [s, 0, 0, c]])
    np.testing.assert_allclose(U("YYPowGate", exponent=1 / 3, global_shift=-0.5), want, atol=1e-12)


def test_zz_matrix_known_answers():
    # ops/parity_gates_test.py:221-235; ZZPowGate(0.25, global_shift=-0.5) is the QAOA cost gate
    # (examples/qaoa.py:128-146: rzz(r) = ZZPowGate(2r/pi, global_shift=-1/2))
    np.testing.assert_allclose(U("ZZPowGate"), np.diag([1, -1, -1, 1]), atol=1e-12)
    np.testing.assert_allclose(U("ZZPowGate", exponent=2.0), np.eye(4), atol=1e-12)
    b = 1j**0.25
    a = np.conj(b)
    np.testing.assert_allclose(U("ZZPowGate", exponent=0.25, global_shift=-0.5), np.diag([a, b, b, a]), atol=1e-12)


# ------------------------------------------------------------------------------ two- and three-qubit gates
@pytest.mark.parametrize(
    "theta,phi,want",
    [
        # ops/fsim_gate_test.py:127-228
        (0, 0, np.eye(4)),
        (np.pi / 2, 0, [[1, 0, 0, 0], [0, 0, -1j, 0], [0, -1j, 0, 0], [0, 0, 0, 1]]),
        (-np.pi / 2, 0, [[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]]),
        (np.pi, 0, np.diag([1, -1, -1, 1])),
        (0, np.pi / 2, np.diag([1, 1, 1, -1j])),
        (0, -np.pi / 2, np.diag([1, 1, 1, 1j])),
        (0, np.pi, np.diag([1, 1, 1, -1])),
        # ops/fsim_gate_test.py:242-253
        (np.pi / 4, np.pi / 3, [[1, 0, 0, 0], [0, np.sqrt(0.5), -1j * np.sqrt(0.5), 0],
                                [0, -1j * np.sqrt(0.5), np.sqrt(0.5), 0], [0, 0, 0, 0.5 - 1j * np.sqrt(0.75)]]),
    ],
)
def test_fsim_known_answers(theta, phi, want):
    np.testing.assert_allclose(U("FSimGate", theta=theta, phi=phi), np.asarray(want, dtype=complex), atol=1e-12)


@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_diagonal_gate_known_answers(n):
    # ops/diagonal_gate_test.py:148-155 (_candidate_angles, :23)
    angles = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53][: 2**n]
    np.testing.assert_allclose(np.diag(U("DiagonalGate", angles=angles)), np.exp(1j * np.array(angles)), atol=1e-12)
    assert np.count_nonzero(U("DiagonalGate", angles=angles) - np.diag(np.diag(U("DiagonalGate", angles=angles)))) == 0


def test_three_qubit_gates_known_answers():
    # ops/three_qubit_gates_test.py:55-130
    ccx = np.eye(8, dtype=complex)
    ccx[6:, 6:] = [[0, 1], [1, 0]]
    np.testing.assert_allclose(U("CCXPowGate"), ccx, atol=1e-12)
    half = np.eye(8, dtype=complex)
    half[6:, 6:] = [[0.5 + 0.5j, 0.5 - 0.5j], [0.5 - 0.5j, 0.5 + 0.5j]]
    np.testing.assert_allclose(U("CCXPowGate", exponent=0.5), half, atol=1e-12)
    np.testing.assert_allclose(U("CCZPowGate"), np.diag([1, 1, 1, 1, 1, 1, 1, -1]), atol=1e-12)
    # CCY = controlled-controlled Y**1 and Y**0.5 (the same tests, :100-128)
    ccy = U("Controlled", sub_family="YPowGate", sub_params={"exponent": 1.0}, control_values=[1, 1])
    want = np.eye(8, dtype=complex)
    want[6:, 6:] = [[0, -1j], [1j, 0]]
    np.testing.assert_allclose(ccy, want, atol=1e-12)
    ccy5 = U("Controlled", sub_family="YPowGate", sub_params={"exponent": 0.5}, control_values=[1, 1])
    want[6:, 6:] = [[0.5 + 0.5j, -0.5 - 0.5j], [0.5 + 0.5j, 0.5 + 0.5j]]
    np.testing.assert_allclose(ccy5, want, atol=1e-12)
    # CSWAP swaps |101> and |110>
    cs = U("CSwapGate")
    assert cs[5, 6] == 1 and cs[6, 5] == 1 and cs[5, 5] == 0 and np.allclose(np.diag(cs)[:5], 1)


def test_controlled_gate_known_answers():
    # ops/controlled_gate_test.py:325-375: CY, C0Y (control value 0), CCH
    cy = U("Controlled", sub_family="Pauli", sub_params={"pauli": "Y"}, control_values=[1])
    np.testing.assert_allclose(cy, [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, -1j], [0, 0, 1j, 0]], atol=1e-12)
    c0y = U("Controlled", sub_family="Pauli", sub_params={"pauli": "Y"}, control_values=[0])
    np.testing.assert_allclose(c0y, [[0, -1j, 0, 0], [1j, 0, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], atol=1e-12)
    cch = U("Controlled", sub_family="HPowGate", sub_params={"exponent": 1.0}, control_values=[1, 1])
    want = np.eye(8, dtype=complex)
    want[6:, 6:] = np.sqrt(0.5) * np.array([[1, 1], [1, -1]])
    np.testing.assert_allclose(cch, want, atol=1e-12)


def test_iswap_family_known_answers():
    # ops/swap_gates_test.py:94-148 (ISWAP, ISWAP_INV, SQRT_ISWAP, SQRT_ISWAP_INV)
    r = 1 / np.sqrt(2)
    np.testing.assert_allclose(U("ISwapPowGate"), [[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]], atol=1e-12)
    np.testing.assert_allclose(U("ISwapPowGate", exponent=-1.0), [[1, 0, 0, 0], [0, 0, -1j, 0], [0, -1j, 0, 0], [0, 0, 0, 1]], atol=1e-12)
    np.testing.assert_allclose(U("ISwapPowGate", exponent=0.5), [[1, 0, 0, 0], [0, r, 1j * r, 0], [0, 1j * r, r, 0], [0, 0, 0, 1]], atol=1e-12)
    np.testing.assert_allclose(U("ISwapPowGate", exponent=-0.5), [[1, 0, 0, 0], [0, r, -1j * r, 0], [0, -1j * r, r, 0], [0, 0, 0, 1]], atol=1e-12)


@pytest.mark.parametrize("z0,y,z1", list(itertools.product([-0.2, 0.25, 1.0], [0.2, 0.5, 1.0], [-0.2, 0.0, 0.5])))
def test_phased_xz_from_zyz_exponents(z0, y, z1):
    # ops/phased_x_z_gate_test.py:52-63 with PhasedXZGate.from_zyz_exponents (ops/phased_x_z_gate.py:85-90):
    # PhasedXZ(a = -z0 + 1/2, x = y, z = z0 + z1) equals Z**z0, Y**y, Z**z1 in time order up to a phase
    got = U("PhasedXZGate", axis_phase_exponent=-z0 + 0.5, x_exponent=y, z_exponent=z0 + z1)
    zyz = U("ZPowGate", exponent=z1) @ U("YPowGate", exponent=y) @ U("ZPowGate", exponent=z0)
    assert allclose_up_to_global_phase(got, zyz)


def test_phased_xz_exact_at_half_integers():
    # ops/phased_x_z_gate_test.py:249-262: entries are multiples of 1/2 (real and imaginary parts)
    for x, z, a in itertools.product([-0.5, 0.0, 0.5, 1.0], repeat=3):
        u = U("PhasedXZGate", x_exponent=x, z_exponent=z, axis_phase_exponent=a)
        np.testing.assert_allclose(u, np.round(u * 2) / 2, atol=1e-12)


# ------------------------------------------------------------------------------ channels (Kraus lists, in order)
def test_asymmetric_depolarize_kraus_order():
    # ops/common_channels_test.py:61-68
    assert_kraus(kraus("asymmetric_depolarize", p_x=0.1, p_y=0.2, p_z=0.3),
                 [np.sqrt(0.4) * I2, np.sqrt(0.1) * X, np.sqrt(0.2) * Y, np.sqrt(0.3) * Z])


def test_two_qubit_depolarize_kraus_order():
    # ops/common_channels_test.py:150-172
    P = [I2, X, Y, Z]
    want = [np.sqrt(0.85) * np.eye(4)] + [np.sqrt(0.01) * np.kron(a, b) for a in P for b in P][1:]
    assert_kraus(kraus("depolarize", p=0.15, n_qubits=2), want)


def test_generalized_amplitude_damp_kraus_order():
    # ops/common_channels_test.py:326-337
    assert_kraus(kraus("generalized_amplitude_damp", p=0.1, gamma=0.3), [
        np.sqrt(0.1) * np.array([[1.0, 0.0], [0.0, np.sqrt(1.0 - 0.3)]]),
        np.sqrt(0.1) * np.array([[0.0, np.sqrt(0.3)], [0.0, 0.0]]),
        np.sqrt(0.9) * np.array([[np.sqrt(1.0 - 0.3), 0.0], [0.0, 1.0]]),
        np.sqrt(0.9) * np.array([[0.0, 0.0], [np.sqrt(0.3), 0.0]]),
    ])


def test_reset_kraus_order():
    # ops/common_channels_test.py:450-455
    assert_kraus(kraus("reset"), [np.array([[1.0, 0.0], [0.0, 0]]), np.array([[0.0, 1.0], [0.0, 0.0]])])


def test_phase_damp_kraus_order():
    # ops/common_channels_test.py:545-553
    assert_kraus(kraus("phase_damp", gamma=0.3),
                 [np.array([[1.0, 0.0], [0.0, np.sqrt(1 - 0.3)]]), np.array([[0.0, 0.0], [0.0, np.sqrt(0.3)]])])


def test_bit_and_phase_flip_kraus_order():
    # ops/common_channels_test.py:611-615 and :677-681
    assert_kraus(kraus("phase_flip", p=0.3), [np.sqrt(1.0 - 0.3) * I2, np.sqrt(0.3) * Z])
    assert_kraus(kraus("bit_flip", p=0.3), [np.sqrt(1.0 - 0.3) * I2, np.sqrt(0.3) * X])


# ------------------------------------------------------------------------------ the gates' own strategies
_STRATEGY_CASES = [
    ("XPowGate", {"exponent": 1.0, "global_shift": 0.0}, 1),
    ("XPowGate", {"exponent": 1.0, "global_shift": -0.5}, 1),
    ("YPowGate", {"exponent": 1.0, "global_shift": 0.0}, 1),
    ("YPowGate", {"exponent": 1.0, "global_shift": 0.3}, 1),
    ("ZPowGate", {"exponent": 0.37, "global_shift": 0.0}, 1),
    ("ZPowGate", {"exponent": -1.3, "global_shift": -0.5}, 1),
    ("Rz", {"rads": 0.77}, 1),
    ("HPowGate", {"exponent": 1.0, "global_shift": 0.0}, 1),
    ("HPowGate", {"exponent": 1.0, "global_shift": 0.25}, 1),
    ("CZPowGate", {"exponent": 1.0, "global_shift": 0.0}, 2),
    ("CZPowGate", {"exponent": 0.41, "global_shift": -0.5}, 2),
    ("CXPowGate", {"exponent": 1.0, "global_shift": 0.0}, 2),
    ("CXPowGate", {"exponent": 1.0, "global_shift": 0.2}, 2),
    ("ZZPowGate", {"exponent": 0.25, "global_shift": -0.5}, 2),
    ("ZZPowGate", {"exponent": 1.7, "global_shift": 0.0}, 2),
]


@pytest.mark.parametrize("family,params,k", _STRATEGY_CASES)
def test_gate_strategies_equal_kron_products(family, params, k):
    """Each gate's own slice strategy (subspace indices little-endian over the target axes) equals
    the dense Kronecker product of its pinned matrix, for target axes in both orders."""
    n = 4
    rng = np.random.RandomState(3)
    psi = (rng.randn(2**n) + 1j * rng.randn(2**n)).astype(np.complex128)
    for axes in ([1] if k == 1 else [1, 3], [3] if k == 1 else [3, 0]):
        t = psi.copy().reshape((2,) * n)
        r = O.gate_apply_unitary(t, family, params, axes)
        assert r is not None
        M = O.gate_unitary(family, params)
        # dense operator on n qubits: permute the targets to the front, apply, permute back
        rest = [a for a in range(n) if a not in axes]
        perm = list(axes) + rest
        v = np.transpose(psi.reshape((2,) * n), perm).reshape(2**k, -1)
        v = (M @ v).reshape((2,) * n)
        want = np.transpose(v, np.argsort(perm)).reshape(-1)
        np.testing.assert_allclose(r.reshape(-1), want, atol=1e-12)


def test_gate_strategies_decline_where_cirq_declines():
    # XPow/YPow/HPow/CXPow return NotImplemented for exponents other than 1
    t = np.zeros((2, 2, 2), dtype=np.complex128)
    for fam in ("XPowGate", "YPowGate", "HPowGate", "CXPowGate"):
        axes = [0] if fam != "CXPowGate" else [0, 1]
        assert O.gate_apply_unitary(t, fam, {"exponent": 0.5, "global_shift": 0.0}, axes) is None
    assert O.gate_apply_unitary(t, "PhasedXPowGate", {"phase_exponent": 0.25, "exponent": 0.5}, [0]) is None
