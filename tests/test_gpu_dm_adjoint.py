"""GPU tests of the density-matrix adjoint gradient (cirq_amd.dm_adjoint, through the C ABI):
against central finite differences of the oracle's complex128 density-matrix simulation, against
the NumPy kernel model of the same checkpointed sweep, and, for noiseless circuits, against the
state-vector adjoint gradient.  Bar: gradient error <= 1e-5 (max-abs), value within 1e-5."""

import sys

import numpy as np
import pytest
import torch

from cirq_amd import generators
from cirq_amd.adjoint import AdjointGradient, ExpectationLayer
from cirq_amd.dm_adjoint import DMAdjointGradient
import cirq_amd as cirq
from tests.helpers import run_dm_adjoint_model
from tests.test_dm_adjoint_host import _observable, _oracle_dm_ev, _oracle_dm_grad, noisy_circuit

pytestmark = pytest.mark.gpu

GRAD_ATOL = 1e-5


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize(
    "n,layers,tiles,invert,scale",
    [(3, 2, (12, 12), False, 1.0), (3, 2, (4, 4), False, 1.0), (5, 2, (6, 5), False, 1.0), (6, 1, (12, 8), False, 1.0),
     (3, 2, (12, 12), True, 1.0), (5, 2, (6, 5), True, 1.0), (6, 2, (12, 8), True, 0.1), (4, 2, (12, 12), True, 10.0)],
)
def test_dm_adjoint_matches_oracle_finite_differences(gpu, n, layers, tiles, invert, scale):
    _log(f"dm adjoint noisy n={n} layers={layers} tiles={tiles} invert_channels={invert} noise x{scale}")
    c, q = noisy_circuit(n, layers, scale=scale)
    psum = _observable(q)
    g = DMAdjointGradient(c, psum, tile_bits=tiles[0], adjoint_tile_bits=tiles[1], invert_channels=invert)
    vals = np.random.RandomState(n + layers).uniform(-np.pi, np.pi, len(g.symbols))
    value, grad = g.value_and_grad(vals)
    res = cirq.ParamResolver(dict(zip(g.symbols, vals)))
    assert abs(value - _oracle_dm_ev(c, list(q), psum, res)) < 1e-5
    ref = _oracle_dm_grad(c, list(q), psum, g.symbols, vals)
    np.testing.assert_allclose(grad, ref, atol=GRAD_ATOL)
    # the same sweep on the NumPy kernel model
    mv, mg, _ = run_dm_adjoint_model(g.prog, vals)
    assert abs(value - mv) < 1e-5
    np.testing.assert_allclose(grad, mg, atol=GRAD_ATOL)


@pytest.mark.parametrize("every", [1, 3, 1000])
def test_dm_adjoint_checkpoint_spacing(gpu, every):
    _log(f"dm adjoint checkpoint_every={every}")
    c, q = noisy_circuit(4, 2)
    psum = _observable(q)
    base = DMAdjointGradient(c, psum, invert_channels=False)
    other = DMAdjointGradient(c, psum, checkpoint_every=every, invert_channels=False)
    vals = np.random.RandomState(7).uniform(-np.pi, np.pi, len(base.symbols))
    v0, g0 = base.value_and_grad(vals)
    v1, g1 = other.value_and_grad(vals)
    assert abs(v0 - v1) < 1e-6
    np.testing.assert_allclose(g0, g1, atol=1e-6)
    # a second call reuses every buffer and gives the same answer
    v2, g2 = base.value_and_grad(vals)
    assert v2 == v0 and np.array_equal(g2, g0)


def test_noiseless_dm_adjoint_matches_state_vector_adjoint(gpu):
    _log("dm adjoint noiseless 8q HEA vs state-vector adjoint")
    circuit, syms, vals = generators.hardware_efficient_ansatz(8, 2, seed=3)
    qubits = sorted(circuit.all_qubits())
    psum = generators.random_pauli_sum(qubits, 8, 3, seed=4)
    v_sv, g_sv = AdjointGradient(circuit, psum).value_and_grad(vals)
    v_dm, g_dm = DMAdjointGradient(circuit, psum).value_and_grad(vals)
    assert abs(v_sv - v_dm) < 1e-5
    np.testing.assert_allclose(g_dm, g_sv, atol=GRAD_ATOL)


def test_dm_expectation_layer_backward(gpu):
    _log("dm expectation layer")
    c, q = noisy_circuit(3, 1)
    psum = _observable(q)
    layer = ExpectationLayer(c, psum, density_matrix=True)
    x = torch.tensor(np.random.RandomState(2).uniform(-1, 1, len(layer.symbols)), dtype=torch.float64,
                     requires_grad=True)
    y = layer(x)
    y.backward()
    value, grad = layer.engine.value_and_grad(x.detach().numpy())
    assert abs(float(y) - value) < 1e-12
    np.testing.assert_allclose(x.grad.numpy(), grad, atol=1e-12)
