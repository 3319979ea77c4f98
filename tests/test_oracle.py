"""Pins the NumPy oracle (oracle/cirq_np.py) against the reference's known answers and mathematics.

Known answers are the literal expectations of the reference's own tests (file:line cited per
test); the rest are closed forms and brute-force Kronecker products chosen so that a dropped
term, a wrong sign or a transposed operand fails.
"""

import itertools
import math

import numpy as np
import pytest

from oracle import cirq_np as O


# ------------------------------------------------------------------ gate matrices
def test_x_y_z_powers_known_answers():
    # cirq-core/cirq/ops/common_gates_test.py:343-376 (test_z/y/x_unitary)
    X = lambda t: O.gate_unitary("XPowGate", {"exponent": t})
    Y = lambda t: O.gate_unitary("YPowGate", {"exponent": t})
    Z = lambda t: O.gate_unitary("ZPowGate", {"exponent": t})
    assert np.allclose(Z(1), [[1, 0], [0, -1]])
    assert np.allclose(Z(0.5), [[1, 0], [0, 1j]])
    assert np.allclose(Z(0), np.eye(2))
    assert np.allclose(Z(-0.5), [[1, 0], [0, -1j]])
    assert np.allclose(Y(1), [[0, -1j], [1j, 0]])
    assert np.allclose(Y(0.5), np.array([[1 + 1j, -1 - 1j], [1 + 1j, 1 + 1j]]) / 2)
    assert np.allclose(Y(-0.5), np.array([[1 - 1j, 1 - 1j], [-1 + 1j, 1 - 1j]]) / 2)
    assert np.allclose(X(1), [[0, 1], [1, 0]])
    assert np.allclose(X(0.5), np.array([[1 + 1j, 1 - 1j], [1 - 1j, 1 + 1j]]) / 2)
    assert np.allclose(X(-0.5), np.array([[1 - 1j, 1 + 1j], [1 + 1j, 1 - 1j]]) / 2)


def test_cz_known_answers():
    # cirq-core/cirq/ops/common_gates_test.py:81-99
    CZ = lambda t: O.gate_unitary("CZPowGate", {"exponent": t})
    assert np.allclose(CZ(1), np.diag([1, 1, 1, -1]))
    assert np.allclose(CZ(0.5), np.diag([1, 1, 1, 1j]))
    assert np.allclose(CZ(0), np.eye(4))
    assert np.allclose(CZ(-0.5), np.diag([1, 1, 1, -1j]))


def test_rotations_closed_forms():
    # Rx/Ry/Rz docstrings (cirq-core/cirq/ops/common_gates.py:304-330,545-555,872-882)
    for th in (0.3, -1.2, 2.9):
        c, s = math.cos(th / 2), math.sin(th / 2)
        assert np.allclose(O.gate_unitary("Rx", {"rads": th}), [[c, -1j * s], [-1j * s, c]])
        assert np.allclose(O.gate_unitary("Ry", {"rads": th}), [[c, -s], [s, c]])
        assert np.allclose(O.gate_unitary("Rz", {"rads": th}), np.diag([np.exp(-0.5j * th), np.exp(0.5j * th)]))


def test_phased_x_closed_form():
    # cirq-core/cirq/ops/phased_x_gate.py:38-44: Z^p X^t Z^-p
    p, t = 0.25, 0.5
    u = O.gate_unitary("PhasedXPowGate", {"phase_exponent": p, "exponent": t})
    z = np.diag([1, np.exp(1j * np.pi * p)])
    x = O.gate_unitary("XPowGate", {"exponent": t})
    assert np.allclose(u, z @ x @ z.conj())
    # sqrt(W) with W = (X+Y)/sqrt(2): its square is W
    w = (np.array([[0, 1], [1, 0]]) + np.array([[0, -1j], [1j, 0]])) / np.sqrt(2)
    assert np.allclose(u @ u, w)


def test_hadamard_and_swap_family():
    H = O.gate_unitary("HPowGate", {"exponent": 1})
    assert np.allclose(H, np.array([[1, 1], [1, -1]]) / np.sqrt(2))
    sw = O.gate_unitary("SwapPowGate", {"exponent": 1})
    assert np.allclose(sw, np.eye(4)[[0, 2, 1, 3]])
    isw = O.gate_unitary("ISwapPowGate", {"exponent": 1})
    assert np.allclose(isw, [[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]])
    cx = O.gate_unitary("CXPowGate", {"exponent": 1})
    assert np.allclose(cx, np.eye(4)[[0, 1, 3, 2]])
    ccx = O.gate_unitary("CCXPowGate", {"exponent": 1})
    assert np.allclose(ccx, np.eye(8)[[0, 1, 2, 3, 4, 5, 7, 6]])


def test_eigen_global_shift_and_unitarity():
    rng = np.random.RandomState(0)
    for fam in O.EIGEN:
        for _ in range(3):
            t, s = rng.uniform(-2, 2), rng.uniform(-1, 1)
            u = O.eigen_unitary(fam, t, s)
            assert np.allclose(u @ u.conj().T, np.eye(u.shape[0]))
            # global shift multiplies by exp(i pi t s)
            assert np.allclose(u, O.eigen_unitary(fam, t, 0.0) * np.exp(1j * np.pi * t * s))


# ------------------------------------------------------------------ channels
def test_depolarize_and_amplitude_damp_kraus_known_answers():
    X, Y, Z = O._PX, O._PY, O._PZ
    # cirq-core/cirq/ops/common_channels_test.py:140-148
    k = O.channel_kraus("depolarize", {"p": 0.3})
    np.testing.assert_almost_equal(k, (np.sqrt(0.7) * np.eye(2), np.sqrt(0.1) * X, np.sqrt(0.1) * Y, np.sqrt(0.1) * Z))
    # cirq-core/cirq/ops/common_channels_test.py:393-403
    k = O.channel_kraus("amplitude_damp", {"gamma": 0.3})
    np.testing.assert_almost_equal(k, ([[1.0, 0.0], [0.0, np.sqrt(0.7)]], [[0.0, np.sqrt(0.3)], [0.0, 0.0]]))


@pytest.mark.parametrize(
    "fam,params",
    [
        ("depolarize", {"p": 0.2}),
        ("depolarize", {"p": 0.2, "n_qubits": 2}),
        ("amplitude_damp", {"gamma": 0.3}),
        ("generalized_amplitude_damp", {"p": 0.4, "gamma": 0.2}),
        ("phase_damp", {"gamma": 0.6}),
        ("bit_flip", {"p": 0.1}),
        ("phase_flip", {"p": 0.3}),
        ("reset", {}),
    ],
)
def test_channels_are_trace_preserving(fam, params):
    ks = O.channel_kraus(fam, params)
    d = ks[0].shape[0]
    assert np.allclose(sum(k.conj().T @ k for k in ks), np.eye(d))


def test_noisy_expectation_known_answers():
    # cirq-core/cirq/sim/density_matrix_simulator_test.py:891-907
    ops = [("gate", "Pauli", {"pauli": "X"}, (0,)), ("channel", "amplitude_damp", {"gamma": 0.1}, (0,))]
    rho = O.simulate_density_matrix(ops, 1)
    assert abs(O.expectation_from_density_matrix(rho, (1, {0: "Z"}), 1) - (-0.8)) < 1e-6
    assert abs(O.expectation_from_density_matrix(rho, (1, {0: "X"}), 1)) < 1e-6
    ops = [("gate", "HPowGate", {"exponent": 1}, (0,)), ("channel", "depolarize", {"p": 0.3}, (0,))]
    rho = O.simulate_density_matrix(ops, 1)
    assert abs(O.expectation_from_density_matrix(rho, (1, {0: "Z"}), 1)) < 1e-6
    assert abs(O.expectation_from_density_matrix(rho, (1, {0: "X"}), 1) - 0.6) < 1e-6


# ------------------------------------------------------------------ expectations
def test_expectation_two_qubit_known_answers():
    # cirq-core/cirq/ops/linear_combinations_test.py:1399-1434
    psum1 = [(1, {0: "Z"}), (3.2, {1: "Z"})]
    psum2 = [(-1, {0: "X"}), (2, {1: "X"})]
    ev = lambda st, ps: sum(O.expectation_from_state_vector(st, t, 2) for t in ps)
    wf1 = np.array([0, 1, 0, 0], dtype=complex)
    assert abs(ev(wf1, psum1) - (-2.2)) < 1e-7
    assert abs(ev(wf1, psum2)) < 1e-7
    wf2 = np.array([1, 1, 1, 1], dtype=complex) / 2
    assert abs(ev(wf2, psum1)) < 1e-7
    assert abs(ev(wf2, psum2) - 1) < 1e-7
    psum3 = [(1, {0: "Z"}), (1, {1: "X"})]
    wf3 = np.array([1, 1, 0, 0], dtype=complex) / np.sqrt(2)
    assert abs(ev(wf3, psum3) - 2) < 1e-7
    psum3_swapped = [(1, {1: "Z"}), (1, {0: "X"})]
    assert abs(ev(wf3, psum3_swapped)) < 1e-7


def test_expectation_matches_dense_operator():
    rng = np.random.RandomState(1)
    n = 4
    psi = rng.randn(16) + 1j * rng.randn(16)
    psi /= np.linalg.norm(psi)
    for _ in range(10):
        paulis = {a: "XYZ"[rng.randint(3)] for a in rng.choice(n, size=rng.randint(1, n + 1), replace=False)}
        mats = [O._PM[paulis.get(a, "I")] for a in range(n)]
        P = mats[0]
        for m in mats[1:]:
            P = np.kron(P, m)
        c = rng.randn() + 1j * rng.randn()
        got = O.expectation_from_state_vector(psi, (c, paulis), n)
        assert abs(got - c * np.vdot(psi, P @ psi)) < 1e-12
        rho = np.outer(psi, psi.conj())
        got = O.expectation_from_density_matrix(rho, (1.0, paulis), n)
        assert abs(got - np.real(np.trace(rho @ P))) < 1e-12


# ------------------------------------------------------------------ application kernels
def _dense_full(u, axes, n):
    """Full 2^n matrix of u on `axes` (big-endian), by permuting a kron product."""
    k = len(axes)
    others = [a for a in range(n) if a not in axes]
    full = np.kron(u, np.eye(2 ** (n - k)))
    perm = list(axes) + others  # tensor axis order of `full`
    t = full.reshape((2,) * (2 * n))
    inv = np.argsort(perm)
    t = t.transpose(list(inv) + [n + i for i in inv])
    return t.reshape(2**n, 2**n)


def test_apply_unitary_matches_kron_bruteforce():
    # the rule of cirq-core/cirq/linalg/transformations_test.py:122-191 (einsum vs kron-dot)
    rng = np.random.RandomState(2)
    n = 5
    for k in (1, 2, 3):
        for _ in range(4):
            axes = list(rng.choice(n, size=k, replace=False))
            a = rng.randn(2**k, 2**k) + 1j * rng.randn(2**k, 2**k)
            u, _ = np.linalg.qr(a)
            psi = rng.randn(2**n) + 1j * rng.randn(2**n)
            got = O.apply_unitary(psi.reshape((2,) * n), u, axes).reshape(-1)
            want = _dense_full(u, axes, n) @ psi
            assert np.allclose(got, want)


def test_apply_channel_matches_superoperator():
    rng = np.random.RandomState(3)
    n = 3
    a = rng.randn(8, 8) + 1j * rng.randn(8, 8)
    rho = a @ a.conj().T
    rho /= np.trace(rho)
    ks = O.channel_kraus("depolarize", {"p": 0.3})
    got = O.apply_channel(rho.reshape((2,) * 6), ks, [1], n).reshape(8, 8)
    want = sum(_dense_full(K, [1], n) @ rho @ _dense_full(K, [1], n).conj().T for K in ks)
    assert np.allclose(got, want)


def test_density_matrix_of_pure_circuit_is_projector():
    # cirq-core/cirq/sim/density_matrix_simulator_test.py:543-556
    ops = [
        ("gate", "HPowGate", {"exponent": 1}, (0,)),
        ("gate", "CXPowGate", {"exponent": 1}, (0, 1)),
        ("gate", "YPowGate", {"exponent": 0.3}, (2,)),
        ("gate", "CZPowGate", {"exponent": 0.7}, (2, 1)),
    ]
    psi = O.simulate_state_vector(ops, 3)
    rho = O.simulate_density_matrix(ops, 3)
    assert np.allclose(rho, np.outer(psi, psi.conj()))


# ------------------------------------------------------------------ sampling
def test_sample_seed_known_answer():
    # cirq-core/cirq/sim/state_vector_test.py:137-150 and density_matrix_utils_test.py:77-93
    state = np.ones(2) / np.sqrt(2)
    want = [[0], [1], [0], [1], [1], [0], [0], [1], [1], [1]]
    s = O.sample_state_vector(state, [0], 10, np.random.RandomState(1234))
    assert np.array_equal(s, want)
    s = O.sample_density_matrix(0.5 * np.eye(2), [0], 10, np.random.RandomState(1234))
    assert np.array_equal(s, want)


def test_sample_big_endian_and_orders():
    # cirq-core/cirq/sim/state_vector_test.py:62-95
    for perm in itertools.permutations([0, 1, 2]):
        for x in range(8):
            st = np.zeros(8, dtype=np.complex64)
            st[x] = 1
            got = O.sample_state_vector(st, list(perm), 1, np.random.RandomState(0))
            assert got.tolist() == [[1 & (x >> (2 - p)) for p in perm]]


def test_terminal_measurement_seed_known_answer():
    # cirq-core/cirq/sim/sparse_simulator_test.py:1046-1110: X**0.5 then measure, seed 1234, 30 reps twice
    psi = O.simulate_state_vector([("gate", "XPowGate", {"exponent": 0.5}, (0,))], 1, dtype=np.complex64)
    prng = np.random.RandomState(1234)
    r1 = O.sample_state_vector(psi, [0], 30, prng)[:, 0]
    r2 = O.sample_state_vector(psi, [0], 30, prng)[:, 0]
    assert "".join(map(str, r1)) == "010110011101110110110100110101"
    assert "".join(map(str, r2)) == "101011010100011101010110111111"


# ------------------------------------------------------------------ NumPy float32 arithmetic model
def _np_float32_sum_model(a):
    """NumPy's float32 sum: a sequential sum over 8192-element buffers of pairwise sums."""

    def pw(x):
        n = len(x)
        if n < 8:
            r = np.float32(0)
            for v in x:
                r = np.float32(r + v)
            return r
        if n <= 128:
            r = [x[j] for j in range(8)]
            i = 8
            while i < n - (n % 8):
                for j in range(8):
                    r[j] = np.float32(r[j] + x[i + j])
                i += 8
            res = np.float32(
                np.float32(np.float32(r[0] + r[1]) + np.float32(r[2] + r[3]))
                + np.float32(np.float32(r[4] + r[5]) + np.float32(r[6] + r[7]))
            )
            while i < n:
                res = np.float32(res + x[i])
                i += 1
            return res
        n2 = n // 2
        n2 -= n2 % 8
        return np.float32(pw(x[:n2]) + pw(x[n2:]))

    tot = np.float32(0)
    for k in range(0, len(a), 8192):
        tot = np.float32(tot + pw(a[k : k + 8192]))
    return tot


@pytest.mark.parametrize("n", [1, 3, 7, 8, 12, 13, 15, 17])
def test_numpy_float32_sum_order(n):
    """The GPU normalisation sum reproduces this order; pin it against NumPy itself."""
    rng = np.random.RandomState(n)
    p = rng.exponential(size=1 << n).astype(np.float32)
    assert _np_float32_sum_model(p) == p.sum()


def test_numpy_complex_abs2_uses_fma():
    """(psi * conj(psi)).real in complex64 evaluates fma(a, a, b*b) on this host; the GPU does the same."""
    rng = np.random.RandomState(0)
    x = (rng.randn(1 << 16) + 1j * rng.randn(1 << 16)).astype(np.complex64)
    p = (x * x.conj()).real
    a = x.real.astype(np.float64)
    b2 = (x.imag * x.imag).astype(np.float32).astype(np.float64)
    fma = (a * a + b2).astype(np.float32)  # exact a*a in float64, one rounding
    assert np.array_equal(p, fma)
