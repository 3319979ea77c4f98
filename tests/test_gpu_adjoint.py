"""GPU tests of the adjoint gradient kernel (through the C ABI) against finite differences of the
oracle and against the GPU parameter-shift rule.  Bar: gradient error <= 1e-5 (max-abs)."""

import sys

import numpy as np
import pytest
import sympy
import torch

import cirq_amd as cirq
from cirq_amd import generators
from cirq_amd.adjoint import AdjointGradient, ExpectationLayer
from tests.test_adjoint_host import _mixed_circuit, _oracle_ev, _oracle_grad

pytestmark = pytest.mark.gpu

GRAD_ATOL = 1e-5


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("tiles", [(5, 5, 4), (3, 3, 1), (2, 3, 0)])
def test_adjoint_mixed_circuit_matches_finite_differences(gpu, tiles):
    _log(f"adjoint mixed circuit tiles={tiles}")
    circuit, syms = _mixed_circuit()
    qubits = sorted(circuit.all_qubits())
    psum = generators.random_pauli_sum(qubits, 6, 3, seed=3)
    L, La, c = tiles
    ag = AdjointGradient(circuit, psum, tile_bits=L, adjoint_tile_bits=La, coalesce_bits=c)
    vals = np.array([0.37, -1.1, 0.52, 0.81])
    value, grad = ag.value_and_grad(vals)
    assert abs(value - _oracle_ev(circuit, qubits, psum, cirq.ParamResolver(dict(zip(syms, vals))))) < 1e-5
    np.testing.assert_allclose(grad, _oracle_grad(circuit, qubits, psum, syms, vals), atol=GRAD_ATOL)
    # the reverse sweep un-computes psi back to |0...0>
    back = ag.psi.cpu().numpy()
    assert abs(back[0] - 1) < 1e-5 and np.abs(back[1:]).max() < 1e-5


@pytest.mark.parametrize("n,layers,tiles", [(8, 2, (6, 5, 4)), (12, 3, (12, 11, 4)), (14, 2, (10, 9, 4))])
def test_adjoint_hea_matches_oracle_and_parameter_shift(gpu, n, layers, tiles):
    _log(f"adjoint HEA n={n} layers={layers} tiles={tiles}")
    circuit, syms, vals = generators.hardware_efficient_ansatz(n, layers, seed=n)
    qubits = sorted(circuit.all_qubits())
    psum = generators.random_pauli_sum(qubits, 8, 4, seed=n)
    L, La, c = tiles
    ag = AdjointGradient(circuit, psum, tile_bits=L, adjoint_tile_bits=La, coalesce_bits=c)
    value, grad = ag.value_and_grad(vals)
    v2, g2 = ag.parameter_shift(vals)
    assert abs(value - v2) < 1e-5
    np.testing.assert_allclose(grad, g2, atol=GRAD_ATOL)
    if n <= 12:
        np.testing.assert_allclose(grad, _oracle_grad(circuit, qubits, psum, syms, vals), atol=GRAD_ATOL)


def test_adjoint_20q_hea_against_parameter_shift(gpu):
    """The 20-qubit, 10-layer (400 symbol) ansatz at the default tiles."""
    _log("adjoint HEA 20q x 10 layers")
    circuit, syms, vals = generators.hardware_efficient_ansatz(20, 10, seed=0)
    qubits = sorted(circuit.all_qubits())
    psum = generators.random_pauli_sum(qubits, 32, 4, seed=0)
    ag = AdjointGradient(circuit, psum)
    assert len(ag.symbols) == 400
    value, grad = ag.value_and_grad(vals)
    v2, g2 = ag.parameter_shift(vals)
    assert abs(value - v2) < 1e-5
    assert np.abs(grad - g2).max() <= GRAD_ATOL


def test_qaoa_gradient(gpu):
    _log("adjoint QAOA 12 nodes p=2")
    circuit, syms, vals, ham, _ = generators.qaoa_maxcut(12, 2, seed=1)
    qubits = sorted(circuit.all_qubits())
    ag = AdjointGradient(circuit, ham)
    value, grad = ag.value_and_grad(dict(zip(syms, vals)))
    want = dict(zip(syms, _oracle_grad(circuit, qubits, ham, syms, vals)))
    np.testing.assert_allclose(grad, [want[s] for s in ag.symbols], atol=GRAD_ATOL)
    # complex64 amplitudes: the value's error scales with the observable's norm (sum |c_t| = 18)
    norm_bound = sum(abs(complex(t.coefficient)) for t in ham)
    assert abs(value - _oracle_ev(circuit, qubits, ham, cirq.ParamResolver(dict(zip(syms, vals))))) < 1e-5 * norm_bound


def test_expectation_layer_autograd(gpu):
    _log("ExpectationLayer autograd")
    circuit, syms, vals = generators.hardware_efficient_ansatz(6, 2, seed=4)
    psum = generators.sum_z(sorted(circuit.all_qubits()))
    layer = ExpectationLayer(circuit, psum)
    x = torch.tensor(vals, dtype=torch.float64, requires_grad=True)
    y = 3.0 * layer(x)
    y.backward()
    _, g = layer.engine.value_and_grad(vals)
    np.testing.assert_allclose(x.grad.numpy(), 3.0 * g, atol=1e-12)
    assert layer.symbols == syms


def test_expectation_layer_batched_rows_and_no_grad(gpu):
    """(B, S) input: one value per row and per-row gradients; without grad only forward passes
    run and give the same values."""
    _log("ExpectationLayer batched / no-grad")
    circuit, syms, vals = generators.hardware_efficient_ansatz(6, 2, seed=4)
    psum = generators.random_pauli_sum(sorted(circuit.all_qubits()), 5, 3, seed=4)
    layer = ExpectationLayer(circuit, psum)
    X = torch.tensor(np.stack([vals, 0.5 * vals, -vals]), dtype=torch.float64, requires_grad=True)
    y = layer(X)
    assert y.shape == (3,)
    w = torch.tensor([1.0, 2.0, 3.0], dtype=torch.float64)
    (w * y).sum().backward()
    for i in range(3):
        v, g = layer.engine.value_and_grad(X[i].detach().numpy())
        assert abs(float(y[i]) - v) < 1e-12
        np.testing.assert_allclose(X.grad[i].numpy(), float(w[i]) * g, atol=1e-12)
    with torch.no_grad():
        y2 = layer(X)
    np.testing.assert_allclose(y2.numpy(), y.detach().numpy(), atol=1e-12)
    assert layer(torch.tensor(vals, dtype=torch.float64)).shape == ()
    with pytest.raises(ValueError):
        layer(torch.zeros(3, 2, len(syms), dtype=torch.float64))


def test_adjoint_rejects_non_unitary_scale(gpu):
    q = cirq.LineQubit.range(2)
    t = sympy.Symbol("t")
    circuit = cirq.Circuit(cirq.rx(t).on(q[0]), cirq.PauliString(qubit_pauli_map={q[1]: cirq.X}, coefficient=2.0))
    with pytest.raises(ValueError):
        AdjointGradient(circuit, cirq.Z(q[0]))


def test_adjoint_large_coefficients_through_mfma_blocks(gpu):
    """phi = H psi with sum |c| = 2000 on 10 qubits, reversed through MFMA blocks: the sweep starts
    from (H / s) psi (s a power of two >= sum |c|) so phi stays in the split-FP16 range."""
    _log("adjoint with large observable coefficients")
    import cirq_amd as cirq_local

    g1 = generators.grid_circuit(2, 5, 12, seed=1)
    qs = sorted(g1.all_qubits())
    syms = sympy.symbols(f"a0:{len(qs)}")
    circuit = cirq_local.Circuit(list(g1.all_operations()) + [cirq_local.rx(syms[i]).on(q) for i, q in enumerate(qs)]
                                 + list(generators.grid_circuit(2, 5, 12, seed=2).all_operations()))
    psum = 200.0 * generators.sum_z(qs)
    ag = AdjointGradient(circuit, psum)
    assert ag.prog.phi_scale == 2048.0
    assert int(((ag.prog.rev.ops_arr["kind"] & 0xFF) == 5).sum()) > 0  # MFMA blocks in the sweep
    vals = np.linspace(-1.0, 1.2, len(qs))
    value, grad = ag.value_and_grad(vals)
    v2, g2 = ag.parameter_shift(vals)
    assert np.all(np.isfinite(grad))
    assert abs(value - v2) < 1e-5 * 2000
    np.testing.assert_allclose(grad, g2, atol=1e-5 * 2000)
