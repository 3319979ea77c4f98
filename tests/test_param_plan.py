"""The compiled parameter plan (tape.ParamPlan) fills a program's matrix buffer exactly as calling
every entry's payload transform does (tape.fill_param_mats_direct), for every parameterised gate
family, controls, derivative payloads of adjoint programs, density-matrix lowerings (whose
unsplit diagonals are not affine and keep the per-entry call) and gates too wide to probe."""

import numpy as np
import pytest
import sympy

import cirq_amd as cirq
from cirq_amd import generators, tape
from cirq_amd.adjoint import AdjointProgram
from cirq_amd.dm_adjoint import DMAdjointProgram


def _mixed_circuit():
    a, b, c = sympy.symbols("a b c")
    q = cirq.LineQubit.range(6)
    ops = [
        cirq.rx(a).on(q[0]), cirq.ry(2 * b + 0.3).on(q[1]), cirq.rz(c / 3).on(q[2]),
        cirq.XPowGate(exponent=a, global_shift=-0.5).on(q[3]), cirq.HPowGate(exponent=b).on(q[4]),
        cirq.ZPowGate(exponent=a * b).on(q[5]),  # not affine in one symbol: evaluated one by one
        cirq.CZPowGate(exponent=c).on(q[0], q[1]), cirq.CXPowGate(exponent=a).on(q[2], q[3]),
        cirq.ZZPowGate(exponent=b, global_shift=-0.5).on(q[4], q[5]), cirq.XXPowGate(exponent=c).on(q[1], q[2]),
        cirq.YYPowGate(exponent=a).on(q[3], q[4]), cirq.ISwapPowGate(exponent=b).on(q[0], q[5]),
        cirq.SwapPowGate(exponent=c).on(q[2], q[4]), cirq.FSimGate(theta=a, phi=b).on(q[1], q[3]),
        cirq.PhasedXPowGate(phase_exponent=b, exponent=c).on(q[0]),
        cirq.PhasedXZGate(x_exponent=a, z_exponent=0.25, axis_phase_exponent=c).on(q[5]),
        cirq.CCZPowGate(exponent=a).on(q[0], q[2], q[4]),
        cirq.ControlledGate(cirq.ry(b), 1).on(q[3], q[1]),
        cirq.ControlledGate(cirq.ZPowGate(exponent=c), 2, control_values=(1, 0)).on(q[0], q[5], q[2]),
        cirq.DiagonalGate([a, b, 0.1, c]).on(q[4], q[1]),
        cirq.rx(a).on(q[0]), cirq.ry(b).on(q[0]),
    ]
    return cirq.Circuit(ops), [a, b, c]


def _points(n_sym, B, seed):
    return np.random.default_rng(seed).uniform(-3, 3, (n_sym, B))


def _check(packed, ev, symvals):
    want = tape.fill_param_mats_direct(packed, ev, symvals)
    got = tape.fill_param_mats(packed, ev, symvals)
    assert got.dtype == want.dtype and got.shape == want.shape
    np.testing.assert_array_equal(got, want)
    return packed._param_plan[1]


@pytest.mark.parametrize("dm", [False, True])
def test_plan_matches_per_entry_transforms(dm):
    circ, syms = _mixed_circuit()
    qs = sorted(circ.all_qubits())
    low = tape.Lowering(qs, dm)
    low.lower_circuit(circ)
    n_bits = 2 * len(qs) if dm else len(qs)
    for kw in (dict(L=4, coalesce=2), dict(L=5, coalesce=2, blocks=True, kmax=4)):
        packed = tape.compile_ops(low.ops, n_bits, **kw)
        ev = tape.ParamEvaluator(low.param_refs)
        plan = _check(packed, ev, _points(len(ev.symbols), 4, 1))
        _check(packed, ev, _points(len(ev.symbols), 1, 2))
        if dm:  # the unsplit density-matrix diagonal d (x) conj(d) is quadratic in d
            assert plan.nonaffine
        else:
            assert not plan.nonaffine and plan.S.nnz > 0


def test_plan_on_adjoint_programs():
    circ, syms = _mixed_circuit()
    psum = cirq.Z(cirq.LineQubit(0)) * cirq.X(cirq.LineQubit(3)) + 0.5 * cirq.Y(cirq.LineQubit(5))
    prog = AdjointProgram(circ, psum, tile_bits=4, adjoint_tile_bits=4)
    sv = _points(len(prog.symbols), 3, 3)
    for packed in (prog.fwd, prog.rev):
        plan = _check(packed, prog.evaluator, sv)
        assert not plan.nonaffine
    hea, hsyms, vals = generators.hardware_efficient_ansatz(8, 3, seed=0)
    prog = AdjointProgram(hea, generators.random_pauli_sum(sorted(hea.all_qubits()), 6, 3, seed=1))
    for packed in (prog.fwd, prog.rev):
        _check(packed, prog.evaluator, prog.symbol_values(vals))


def test_plan_on_density_matrix_adjoint_program():
    noisy, syms, vals = generators.noisy_line_ansatz(4, 2, seed=2, p=0.05, gamma=0.05)
    obs = sum(cirq.Z(q) for q in sorted(noisy.all_qubits()))
    for kw in ({}, {"invert_channels": False}):
        prog = DMAdjointProgram(noisy, obs, tile_bits=5, adjoint_tile_bits=5, **kw)
        sv = prog.symbol_values(vals)
        for packed in (prog.fwd, prog.rev):
            _check(packed, prog.evaluator, sv)


def test_plan_on_wide_gate():
    # a 4-qubit parameterised diagonal (16 x 16 gate matrix) is not probed: per-entry call
    t, u = sympy.symbols("t u")
    q = cirq.LineQubit.range(8)
    circ = cirq.Circuit(cirq.DiagonalGate([t, u] * 8).on(*q[:4]), cirq.rx(t).on(q[6]))
    low = tape.Lowering(q, False)
    low.lower_circuit(circ)
    packed = tape.compile_ops(low.ops, 8, L=6, coalesce=2)
    ev = tape.ParamEvaluator(low.param_refs)
    plan = _check(packed, ev, _points(2, 2, 4))
    assert len(plan.nonaffine) == 1 and plan.S.nnz > 0
