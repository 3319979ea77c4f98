"""Straight-line pass kernels (cirq_amd.specialize) against the op interpreter and the oracle.

The specialised kernels call the interpreter's device functions in the same order with the bit
positions as constants, so their states equal the interpreter's up to the compiler's choice of
fused multiply-adds (max-abs <= 1e-6 here, 1.2e-10 measured at 30 qubits); against the complex128
oracle the simulator's bars hold (max-abs <= 1e-5, fidelity >= 1 - 1e-6).  The programs cover
register-bit dense ops with and without local controls, outer-bit matrix variants and outer
controls, diagonals, pair phases, PERMUTE re-layouts, LDS dense gates and MFMA blocks, on state
vectors and density matrices.
"""

import sys

import numpy as np
import pytest
import torch

import cirq_amd as cirq
from cirq_amd import engine, generators, tape
from oracle import cirq_np as O
from tests.helpers import fidelity, to_oracle_ops
from tests.test_host import _random_circuit

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _both(packed, device):
    """The program's result from |0..0> through the interpreter and through specialised kernels."""
    n = packed.n_bits
    dp = engine.DeviceProgram(packed, device)
    a = engine.new_state(n, device)
    dp.run(a, basis=0)
    b = engine.new_state(n, device)
    dp.specialize().run(b, basis=0)
    torch.cuda.synchronize()
    return a, b, dp


@pytest.mark.parametrize("seed", range(3))
def test_random_programs_match_interpreter_and_oracle(gpu, seed):
    _log(f"specialised random circuit seed={seed}")
    rng = np.random.RandomState(100 + seed)
    n = 12
    c, qs = _random_circuit(n, 160, rng)
    low = tape.Lowering(qs, False)
    low.lower_circuit(c)
    packed = tape.compile_ops(low.ops, n, L=10, coalesce=3, blocks=True, kmax=4)
    a, b, _ = _both(packed, gpu)
    assert float((a - b).abs().max()) <= 1e-6
    want = O.simulate_state_vector(to_oracle_ops(c, qs), n)
    got = b.cpu().numpy()
    assert np.max(np.abs(got - want)) <= 1e-5
    assert fidelity(got, want) >= 1 - 1e-6


@pytest.mark.parametrize("L", [12, 13])
def test_grid_program_with_variants_and_blocks(gpu, L):
    """The benchmark's compilation (3 variant bits, 4-qubit blocks) of a 16-qubit grid circuit."""
    _log(f"specialised grid program L={L}")
    c = generators.grid_circuit(4, 4, 16, seed=2)
    qs = sorted(c.all_qubits())
    low = tape.Lowering(qs, False)
    low.lower_circuit(c)
    packed = tape.compile_ops(low.ops, 16, L=L, coalesce=4, blocks=True, kmax=4, variant_bits=3)
    a, b, dp = _both(packed, gpu)
    assert float((a - b).abs().max()) <= 1e-6
    want = O.simulate_state_vector(to_oracle_ops(c, qs), 16)
    assert fidelity(b.cpu().numpy(), want) >= 1 - 1e-6
    # a run from a state in memory (no basis) gives the interpreter's answer too
    x = b.clone()
    dp.run(x)
    dp.specialized = None
    y = b.clone()
    dp.run(y)
    assert float((x - y).abs().max()) <= 1e-6


def test_simulator_option_and_density_matrix(gpu):
    _log("specialised simulators")
    c = generators.grid_circuit(3, 4, 10, seed=5)
    qs = sorted(c.all_qubits())
    sim = cirq.Simulator(specialize_passes=True, tile_bits=10)
    got = sim.simulate(c).final_state_vector
    want = O.simulate_state_vector(to_oracle_ops(c, qs), 12)
    assert fidelity(got, want) >= 1 - 1e-6 and np.max(np.abs(got - want)) <= 1e-5
    comp = sim._compile(c, qs)
    assert comp.dev.specialized is not None
    q = cirq.LineQubit.range(5)
    noisy = cirq.Circuit(
        [cirq.H(q[i]) for i in range(5)]
        + [cirq.CZ(q[0], q[1]), cirq.rx(0.3).on(q[2]), cirq.depolarize(0.05).on(q[3]),
           cirq.amplitude_damp(0.1).on(q[4]), cirq.CNOT(q[2], q[4]), cirq.rz(0.7).on(q[1])])
    dsim = cirq.DensityMatrixSimulator(specialize_passes=True, tile_bits=8)
    rho = dsim.simulate(noisy).final_density_matrix
    want = O.simulate_density_matrix(to_oracle_ops(noisy, q), 5)
    assert np.max(np.abs(rho - want)) <= 1e-5


@pytest.mark.parametrize("kind", ["hea", "qaoa"])
def test_specialised_gradients_match_interpreter(gpu, kind):
    """AdjointGradient.specialize(): the forward and reverse programs as straight-line kernels give
    the interpreter's value and gradient (register GRAD ops, LDS GRAD ops, diagonals, MFMA blocks
    of the fixed gates, outer-bit variants)."""
    from cirq_amd.adjoint import AdjointGradient

    _log(f"specialised gradients {kind}")
    if kind == "hea":
        circuit, syms, vals = generators.hardware_efficient_ansatz(14, 4, seed=3)
        obs = generators.random_pauli_sum(sorted(circuit.all_qubits()), 8, 3, seed=3)
        kw = dict(tile_bits=10, adjoint_tile_bits=9)
    else:
        circuit, syms, vals, obs, edges = generators.qaoa_maxcut(12, 2, seed=1)
        kw = {}
    ag = AdjointGradient(circuit, obs, **kw)
    v0, g0 = ag.value_and_grad(vals)
    ag.specialize()
    v1, g1 = ag.value_and_grad(vals)
    assert abs(v1 - v0) <= 1e-6 * max(1.0, abs(v0))
    np.testing.assert_allclose(g1, g0, atol=1e-6 * max(1.0, np.abs(g0).max()))


@pytest.mark.parametrize("invert", [True, False])
def test_specialised_density_matrix_gradients_match_interpreter(gpu, invert):
    """DMAdjointGradient.specialize(): forward and reverse pass ranges (steps between checkpoints)
    launched through specialised kernels give the interpreter's value and gradient."""
    from cirq_amd.dm_adjoint import DMAdjointGradient

    _log(f"specialised density-matrix gradients invert={invert}")
    circuit, syms, vals = generators.noisy_line_ansatz(5, 3, seed=2, p=0.02, gamma=0.03)
    H = generators.sum_z(sorted(circuit.all_qubits()))
    g = DMAdjointGradient(circuit, H, invert_channels=invert)
    v0, g0 = g.value_and_grad(vals)
    g.specialize()
    v1, g1 = g.value_and_grad(vals)
    assert abs(v1 - v0) <= 1e-6
    np.testing.assert_allclose(g1, g0, atol=1e-6)


@pytest.mark.parametrize("rows,cols,L,basis", [(4, 4, 11, 0), (4, 4, 11, 0b1011001110100101), (4, 4, 11, 2**16 - 1),
                                                (4, 5, 8, 0), (4, 5, 8, 0b10110011101001011010)])
def test_pass_chain_from_basis_skips_zero_tiles(gpu, rows, cols, L, basis):
    """From a basis state the passes skip the tiles the state cannot reach yet (written as zeros,
    not read) and leave unwritten the tiles the next pass does not read; on buffers filled with NaN
    the result equals the oracle's from the same basis state, through the interpreter and through
    specialised kernels (with 8-bit tiles at 20 qubits the first pass leaves most tiles unwritten)."""
    n = rows * cols
    _log(f"pass chain from basis {basis}, {n} qubits, L={L}")
    c = generators.grid_circuit(rows, cols, 12, seed=4)
    qs = sorted(c.all_qubits())
    low = tape.Lowering(qs, False)
    low.lower_circuit(c)
    packed = tape.compile_ops(low.ops, n, L=L, coalesce=3, blocks=True, kmax=4, variant_bits=2)
    want = O.simulate_state_vector(to_oracle_ops(c, qs), n, initial_state=basis)
    dp = engine.DeviceProgram(packed, gpu)
    a = torch.full((1 << n,), float("nan"), dtype=torch.complex64, device=gpu)
    dp.run(a, basis=basis)
    dp.specialize()
    b = torch.full((1 << n,), float("nan"), dtype=torch.complex64, device=gpu)
    dp.run(b, basis=basis)
    torch.cuda.synchronize()
    for got in (a.cpu().numpy(), b.cpu().numpy()):
        assert not np.isnan(got).any()
        assert np.max(np.abs(got - want)) <= 1e-5
        assert fidelity(got, want) >= 1 - 1e-6
