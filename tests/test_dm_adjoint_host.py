"""CPU tests of the density-matrix adjoint gradient (cirq_amd.dm_adjoint): the step split at
channels, the forward and reversed programs, sqrt checkpointing and the chain rule, executed on
the NumPy kernel model.

Pins: Tr(H rho) and central finite differences of the oracle's complex128 density-matrix
simulation (cirq-core/cirq/sim/density_matrix_simulator.py semantics, restated in oracle/), and
for a noiseless circuit the state-vector adjoint gradient of the same circuit.
"""

import numpy as np
import pytest
import sympy

import cirq_amd as cirq
from cirq_amd.adjoint import AdjointProgram
from cirq_amd.dm_adjoint import DMAdjointProgram
from oracle import cirq_np as O
from tests.helpers import run_adjoint_model, run_dm_adjoint_model, run_pass_model, terms_of, to_oracle_ops

GRAD_ATOL = 2e-5  # complex64 matrices, complex128 model arithmetic


def _oracle_dm_ev(circuit, qubits, psum, resolver):
    ops = to_oracle_ops(circuit, qubits, resolver)
    n = len(qubits)
    rho = O.simulate_density_matrix(ops, n)
    return sum(O.expectation_from_density_matrix(rho, t, n) for t in terms_of(psum, qubits)).real


def _oracle_dm_grad(circuit, qubits, psum, symbols, values, h=1e-5):
    g = np.zeros(len(symbols))
    for i in range(len(symbols)):
        up, dn = dict(zip(symbols, values)), dict(zip(symbols, values))
        up[symbols[i]] = values[i] + h
        dn[symbols[i]] = values[i] - h
        g[i] = (_oracle_dm_ev(circuit, qubits, psum, cirq.ParamResolver(up))
                - _oracle_dm_ev(circuit, qubits, psum, cirq.ParamResolver(dn))) / (2 * h)
    return g


def noisy_circuit(n=3, layers=2, scale=1.0):
    """Rz / Ry rotations with depolarizing noise after each, a parameterised ZZ and a controlled
    rotation on neighbours, amplitude damping on every qubit per layer (diagonal, dense,
    controlled and two-qubit parameterised ops between channels).  `scale` multiplies every
    noise strength."""
    q = cirq.LineQubit.range(n)
    ops = []
    for l in range(layers):
        for i in range(n):
            a, b = sympy.Symbol(f"a{l}_{i}"), sympy.Symbol(f"b{l}_{i}")
            ops += [cirq.rz(a).on(q[i]), cirq.depolarize(0.02 * scale).on(q[i]), cirq.ry(b).on(q[i]),
                    cirq.depolarize(0.01 * scale).on(q[i])]
        for i in range(l % 2, n - 1, 2):
            ops += [cirq.CZ(q[i], q[i + 1]), cirq.depolarize(min(0.03 * scale, 0.9), n_qubits=2).on(q[i], q[i + 1])]
        ops.append(cirq.ZZPowGate(exponent=sympy.Symbol(f"z{l}")).on(q[0], q[n - 1]))
        ops.append(cirq.ControlledGate(cirq.rx(sympy.Symbol(f"c{l}"))).on(q[1], q[0]))
        ops += [cirq.amplitude_damp(min(0.05 * scale, 1.0)).on(x) for x in q]
    return cirq.Circuit(ops), q


def _observable(q):
    return (sum(cirq.Z(x) for x in q) + 0.7 * cirq.X(q[0]) * cirq.X(q[1])
            - 0.4 * cirq.Y(q[1]) * cirq.Z(q[2]) + 0.25 * cirq.Y(q[0]) * cirq.Y(q[2]))


def _values(prog, seed):
    return np.random.RandomState(seed).uniform(-np.pi, np.pi, len(prog.symbols))


def _n_channels(c):
    return sum(1 for op in c.all_operations() if hasattr(op.gate, "_kraus_") or hasattr(op.gate, "_mixture_"))


def test_steps_split_at_channels():
    c, q = noisy_circuit(3, 2)
    prog = DMAdjointProgram(c, _observable(q), tile_bits=4, adjoint_tile_bits=4, invert_channels=False)
    n_chan = _n_channels(c)
    assert prog.n_steps == n_chan + 1 and not prog.undone
    assert prog.steps[0][0] is None
    assert all(ch is not None for ch, _ in prog.steps[1:])
    # every row and column op of a parameterised gate gets its own slot
    assert prog.n_slots == 2 * len(prog.symbols)
    segs = prog.segments()
    assert segs[0][0] == 0 and segs[-1][1] == prog.n_steps
    assert all(b - a <= prog.checkpoint_every for a, b in segs)


def test_well_conditioned_channels_are_undone_in_place():
    """Weak noise: every channel is undone (one step, no checkpoints); stronger noise: the growth
    bound ends steps, and each stretch between step ends stays within it."""
    c, q = noisy_circuit(3, 2, scale=0.1)
    prog = DMAdjointProgram(c, _observable(q))
    assert prog.n_steps == 1 and len(prog.undone) == _n_channels(c)
    c, q = noisy_circuit(3, 2)
    prog = DMAdjointProgram(c, _observable(q), max_amplification=1.2)
    assert 1 < prog.n_steps < _n_channels(c) + 1 and prog.undone
    c, q = noisy_circuit(3, 1, scale=40.0)  # depolarize(0.8): ||S^-1|| = 15 alone exceeds the bound
    prog = DMAdjointProgram(c, _observable(q))
    assert prog.n_steps > 1


@pytest.mark.parametrize("invert,scale", [(False, 1.0), (True, 1.0), (True, 10.0)])
@pytest.mark.parametrize("tiles", [(4, 4), (6, 5), (5, 4), (4, 6)])
def test_noisy_gradient_vs_oracle_finite_differences(tiles, invert, scale):
    c, q = noisy_circuit(3, 2, scale=scale)
    psum = _observable(q)
    prog = DMAdjointProgram(c, psum, tile_bits=tiles[0], adjoint_tile_bits=tiles[1], invert_channels=invert)
    vals = _values(prog, 1)
    value, grad, _ = run_dm_adjoint_model(prog, vals)
    res = cirq.ParamResolver(dict(zip(prog.symbols, vals)))
    assert abs(value - _oracle_dm_ev(c, list(q), psum, res)) < 1e-5
    ref = _oracle_dm_grad(c, list(q), psum, prog.symbols, vals)
    assert np.max(np.abs(grad - ref)) < GRAD_ATOL, np.max(np.abs(grad - ref))
    assert np.max(np.abs(ref)) > 0.05  # a gradient that is not trivially zero


@pytest.mark.parametrize("every", [1, 2, 5, 1000])
def test_checkpoint_spacing_does_not_change_the_result(every):
    c, q = noisy_circuit(3, 2)
    psum = _observable(q)
    base = DMAdjointProgram(c, psum, tile_bits=4, adjoint_tile_bits=4, invert_channels=False)
    prog = DMAdjointProgram(c, psum, tile_bits=4, adjoint_tile_bits=4, checkpoint_every=every, invert_channels=False)
    vals = _values(base, 2)
    v0, g0, _ = run_dm_adjoint_model(base, vals)
    v1, g1, _ = run_dm_adjoint_model(prog, vals)
    assert abs(v0 - v1) < 1e-9 and np.max(np.abs(g0 - g1)) < 1e-9


def test_noiseless_matches_state_vector_adjoint():
    """Without channels Tr(H |psi><psi|) = <psi|H|psi>: the density-matrix sweep (one step, no
    checkpoints) and the state-vector sweep give the same gradient."""
    a, b, t = sympy.symbols("a b t")
    q = cirq.LineQubit.range(3)
    c = cirq.Circuit(
        [cirq.H(x) for x in q],
        cirq.rx(a).on(q[0]), cirq.ry(2 * b + 0.3).on(q[1]), cirq.rz(a * b).on(q[2]),
        cirq.ZZPowGate(exponent=t).on(q[0], q[2]), cirq.FSimGate(theta=a, phi=t).on(q[1], q[2]),
        cirq.ControlledGate(cirq.ry(t)).on(q[0], q[1]), cirq.PhasedXPowGate(phase_exponent=b, exponent=a).on(q[2]),
    )
    psum = _observable(q)
    dm = DMAdjointProgram(c, psum, tile_bits=4, adjoint_tile_bits=4)
    assert dm.n_steps == 1
    sv = AdjointProgram(c, psum, tile_bits=3, adjoint_tile_bits=3)
    assert [s.name for s in dm.symbols] == [s.name for s in sv.symbols]
    vals = np.array([0.4, -1.1, 0.8])
    v_dm, g_dm, _ = run_dm_adjoint_model(dm, vals)
    symvals = sv.symbol_values(vals)
    mf, mr = sv.mats(symvals)
    psi0 = np.zeros(8, dtype=np.complex128)
    psi0[0] = 1
    psi = run_pass_model(sv.fwd, psi0, mf)
    from tests.test_adjoint_host import _apply_psum

    phi = _apply_psum(psi, psum, list(q))
    _, _, gs = run_adjoint_model(sv.rev, psi, phi, mr, sv.n_slots)
    g_sv = sv.chain(gs, symvals)
    assert abs(v_dm - np.vdot(psi, phi).real) < 1e-5
    assert np.max(np.abs(g_dm - g_sv)) < GRAD_ATOL


def test_rejections():
    q = cirq.LineQubit.range(2)
    a = sympy.Symbol("a")
    with pytest.raises(ValueError):
        DMAdjointProgram(cirq.Circuit(cirq.rx(a).on(q[0]), cirq.measure(*q)), cirq.Z(q[0]))
    with pytest.raises(KeyError):
        DMAdjointProgram(cirq.Circuit(cirq.rx(a).on(q[0])), cirq.Z(cirq.LineQubit(7)))
