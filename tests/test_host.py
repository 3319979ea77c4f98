"""CPU tests of the host side: gate semantics, lowering, fusion, scheduling and packing.

The packed programs are executed by the NumPy model of the pass kernel (tests/helpers.py) and
compared with the oracle, so the compiler is checked without a GPU.
"""

import re
from pathlib import Path

import numpy as np
import pytest
import sympy

import cirq_amd as cirq
from cirq_amd import generators, tape
from cirq_amd.simulator import product_state_factors
from oracle import cirq_np as O
from tests.helpers import fidelity, run_pass_model, terms_of, to_oracle_ops

ROOT = Path(__file__).resolve().parents[1]


def _random_circuit(n, depth, rng, with_channels=False):
    qs = cirq.LineQubit.range(n)
    c = cirq.Circuit()
    one = [
        lambda: cirq.X ** rng.uniform(-1, 1),
        lambda: cirq.Y ** rng.uniform(-1, 1),
        lambda: cirq.Z ** rng.uniform(-1, 1),
        lambda: cirq.H,
        lambda: cirq.rx(rng.uniform(-3, 3)),
        lambda: cirq.ry(rng.uniform(-3, 3)),
        lambda: cirq.rz(rng.uniform(-3, 3)),
        lambda: cirq.PhasedXPowGate(phase_exponent=rng.uniform(-1, 1), exponent=rng.uniform(-1, 1)),
        lambda: cirq.PhasedXZGate(
            x_exponent=rng.uniform(-1, 1), z_exponent=rng.uniform(-1, 1), axis_phase_exponent=rng.uniform(-1, 1)
        ),
        lambda: cirq.XPowGate(exponent=rng.uniform(-1, 1), global_shift=rng.uniform(-1, 1)),
    ]
    two = [
        lambda: cirq.CZ ** rng.uniform(-1, 1),
        lambda: cirq.CNOT,
        lambda: cirq.SWAP ** rng.uniform(-1, 1),
        lambda: cirq.ISWAP ** rng.uniform(-1, 1),
        lambda: cirq.XX ** rng.uniform(-1, 1),
        lambda: cirq.YY ** rng.uniform(-1, 1),
        lambda: cirq.ZZPowGate(exponent=rng.uniform(-1, 1), global_shift=-0.5),
        lambda: cirq.FSimGate(rng.uniform(-2, 2), rng.uniform(-2, 2)),
        lambda: cirq.CZ,
    ]
    three = [lambda: cirq.CCZ ** rng.uniform(-1, 1), lambda: cirq.TOFFOLI, lambda: cirq.CSWAP]
    chans = [
        lambda: cirq.depolarize(rng.uniform(0, 0.2)),
        lambda: cirq.amplitude_damp(rng.uniform(0, 0.3)),
        lambda: cirq.phase_damp(rng.uniform(0, 0.3)),
        lambda: cirq.bit_flip(rng.uniform(0, 0.3)),
    ]
    for _ in range(depth):
        r = rng.uniform()
        if r < 0.45:
            q = qs[rng.randint(n)]
            c.append(one[rng.randint(len(one))]().on(q))
        elif r < 0.8:
            a, b = rng.choice(n, 2, replace=False)
            c.append(two[rng.randint(len(two))]().on(qs[a], qs[b]))
        elif r < 0.9:
            a, b, d = rng.choice(n, 3, replace=False)
            c.append(three[rng.randint(len(three))]().on(qs[a], qs[b], qs[d]))
        elif r < 0.95:
            a, b, d = rng.choice(n, 3, replace=False)
            g = (cirq.X ** rng.uniform(-1, 1)).controlled(2, control_values=(int(rng.randint(2)), 1))
            c.append(g.on(qs[a], qs[b], qs[d]))
        elif with_channels:
            c.append(chans[rng.randint(len(chans))]().on(qs[rng.randint(n)]))
        else:
            c.append(cirq.MatrixGate(np.linalg.qr(rng.randn(4, 4) + 1j * rng.randn(4, 4))[0]).on(
                *[qs[i] for i in rng.choice(n, 2, replace=False)]))
    return c, qs


def _compile_sv(c, qs, L, coalesce, fuse=True):
    low = tape.Lowering(qs, False)
    low.lower_circuit(c)
    return tape.compile_ops(low.ops, len(qs), L=L, coalesce=coalesce, fuse=fuse)


# ------------------------------------------------------------------ gate semantics vs oracle
def test_gate_matrices_match_oracle_tables():
    rng = np.random.RandomState(0)
    for _ in range(10):
        c, qs = _random_circuit(4, 30, rng)
        for op, oop in zip(c.all_operations(), to_oracle_ops(c, qs)):
            u = op.gate._unitary_() if op.gate is not None else op._unitary_()
            assert np.allclose(u, O.gate_unitary(oop[1], oop[2]), atol=1e-12), op


def test_pauli_algebra():
    a, b = cirq.LineQubit.range(2)
    assert cirq.X(a) * cirq.Y(a) == cirq.PauliString(qubit_pauli_map={a: cirq.Z}, coefficient=1j)
    assert cirq.Y(a) * cirq.X(a) == cirq.PauliString(qubit_pauli_map={a: cirq.Z}, coefficient=-1j)
    ps = 0.5 * cirq.Z(a) * cirq.Z(b) + cirq.X(b) - cirq.X(b)
    assert len(ps) == 1
    s = cirq.Z(a) + 3.2 * cirq.Z(b)
    assert {complex(t.coefficient) for t in s} == {1, 3.2}
    assert (cirq.X**0.5).exponent == 0.5


# ------------------------------------------------------------------ compiler through the pass model
@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("L,coalesce", [(3, 1), (4, 2), (5, 3), (7, 4)])
def test_pass_model_state_vector(seed, L, coalesce):
    rng = np.random.RandomState(seed)
    n = 7
    c, qs = _random_circuit(n, 60, rng)
    packed = _compile_sv(c, qs, L, coalesce, fuse=seed % 2 == 0)
    psi0 = np.zeros(2**n, dtype=np.complex128)
    psi0[0] = 1
    got = run_pass_model(packed, psi0)
    want = O.simulate_state_vector(to_oracle_ops(c, qs), n)
    assert np.max(np.abs(got - want)) < 2e-6  # matrices are uploaded as complex64


@pytest.mark.parametrize("seed", range(6))
def test_pass_model_density_matrix(seed):
    rng = np.random.RandomState(100 + seed)
    n = 4
    c, qs = _random_circuit(n, 40, rng, with_channels=True)
    low = tape.Lowering(qs, True)
    low.lower_circuit(c)
    packed = tape.compile_ops(low.ops, 2 * n, L=5, coalesce=2)
    rho0 = np.zeros(4**n, dtype=np.complex128)
    rho0[0] = 1
    got = run_pass_model(packed, rho0).reshape(2**n, 2**n)
    want = O.simulate_density_matrix(to_oracle_ops(c, qs), n)
    assert np.max(np.abs(got - want)) < 2e-6  # matrices are uploaded as complex64


def test_pass_model_parametric_sweep():
    rng = np.random.RandomState(7)
    circ, syms, vals = generators.hardware_efficient_ansatz(5, 3, seed=1)
    qs = sorted(circ.all_qubits())
    low = tape.Lowering(qs, False)
    low.lower_circuit(circ)
    packed = tape.compile_ops(low.ops, 5, L=3, coalesce=1)
    ev = tape.ParamEvaluator(low.param_refs)
    resolvers = [cirq.ParamResolver(dict(zip(syms, rng.uniform(-3, 3, len(syms))))) for _ in range(3)]
    mats = tape.fill_param_mats(packed, ev, ev.symbol_values(resolvers))
    for r, m in zip(resolvers, mats):
        psi0 = np.zeros(32, dtype=np.complex128)
        psi0[0] = 1
        got = run_pass_model(packed, psi0, m)
        want = O.simulate_state_vector(to_oracle_ops(circ, qs, resolver=r), 5)
        assert fidelity(got, want) > 1 - 1e-6
        assert np.max(np.abs(got - want)) < 1e-6  # complex64 matrices


def test_pass_model_grid_circuit_and_dm_noise():
    c = generators.grid_circuit(3, 3, 6, seed=3)
    qs = sorted(c.all_qubits())
    packed = _compile_sv(c, qs, 5, 2)
    psi0 = np.zeros(2**9, dtype=np.complex128)
    psi0[0] = 1
    got = run_pass_model(packed, psi0)
    want = O.simulate_state_vector(to_oracle_ops(c, qs), 9)
    assert np.max(np.abs(got - want)) < 2e-6  # matrices are uploaded as complex64
    nc = generators.noisy_line_circuit(3, 3, seed=2, p=0.05, gamma=0.05)
    qs = sorted(nc.all_qubits())
    low = tape.Lowering(qs, True)
    low.lower_circuit(nc)
    packed = tape.compile_ops(low.ops, 6, L=4, coalesce=2)
    rho0 = np.zeros(64, dtype=np.complex128)
    rho0[0] = 1
    got = run_pass_model(packed, rho0).reshape(8, 8)
    want = O.simulate_density_matrix(to_oracle_ops(nc, qs), 3)
    assert np.max(np.abs(got - want)) < 2e-6  # matrices are uploaded as complex64


def test_schedule_respects_tile_and_coalescing():
    c = generators.grid_circuit(4, 5, 10, seed=0)
    qs = sorted(c.all_qubits())
    packed = _compile_sv(c, qs, 10, 4)
    for ps in packed.pass_arr:
        lp = list(ps["lpos"][: ps["L"]])
        assert len(set(lp)) == len(lp) and lp[:4] == [0, 1, 2, 3]
    # fewer passes than gates: the point of the tile scheduler
    n_gates = sum(1 for _ in c.all_operations())
    assert packed.n_passes < n_gates / 10


# ------------------------------------------------------------------ generator (restated structure checks)
def test_random_grid_circuit_structure():
    # cirq-core/cirq/experiments/random_quantum_circuit_generation_test.py:369-404
    qubits = cirq.GridQubit.rect(4, 5)
    c = generators.random_rotations_between_grid_interaction_layers_circuit(
        qubits, 20, pattern=generators.HALF_GRID_STAGGERED_PATTERN, add_final_single_qubit_layer=False, seed=0
    )
    moments = list(c)
    assert len(moments) == 40
    for i in range(0, 40, 2):
        m1 = moments[i]
        assert len(m1) == 20 and all(op.gate in (generators.SQRT_X, generators.SQRT_Y, generators.SQRT_W) for op in m1)
        if i:
            prev = moments[i - 2]
            for q in qubits:
                assert m1.operation_at(q).gate is not prev.operation_at(q).gate
        m2 = moments[i + 1]
        assert all(isinstance(op.gate, cirq.CZPowGate) for op in m2)
        for op in m2:
            a, b = op.qubits
            assert a.is_adjacent(b)
    # CZ counts per layer type A,B,C,D on a 4x5 grid (8/7/8/8)
    assert [len(moments[2 * k + 1]) for k in range(4)] == [8, 7, 8, 8]


def test_product_state_factor_order():
    a, b, c, d = cirq.LineQubit.range(4)
    circ = cirq.Circuit(cirq.H(a), cirq.CNOT(c, a), cirq.X(b), cirq.I(d))
    f = product_state_factors(circ, (a, b, c, d))
    # initial map is over reversed qubits: d, c, b, a; c merges a in (c, a) order
    assert f == [[d], [c, a], [b]]
    circ2 = cirq.Circuit(cirq.CNOT(c, a), cirq.SWAP(a, b))
    f2 = product_state_factors(circ2, (a, b, c, d))
    assert f2 == [[d], [c, b], [a]]


# ------------------------------------------------------------------ study
def test_sweeps_and_resolvers():
    s = cirq.Points("a", [1, 2]) * cirq.Linspace("b", 0, 1, 3)
    rs = list(cirq.to_resolvers(s))
    assert len(rs) == 6 and rs[-1].value_of("b") == 1.0
    z = cirq.Zip(cirq.Points("a", [1, 2]), cirq.Points("b", [3, 4]))
    assert [r.value_of("b") for r in z] == [3, 4]
    x = sympy.Symbol("x")
    r = cirq.ParamResolver({"x": 0.5})
    assert r.value_of(2 * x + 1) == 2.0
    g = cirq.rx(x)._resolve_parameters_(r)
    assert np.allclose(g._unitary_(), cirq.rx(0.5)._unitary_())


# ------------------------------------------------------------------ the C ABI library
def test_library_exports_every_declared_symbol():
    import ctypes

    from cirq_amd import _lib

    header = (ROOT / "include" / "cirq_amd.h").read_text()
    declared = set(re.findall(r"^\s*(?:int|size_t|const char\*)\s+(cqs_\w+)\(", header, re.M))
    assert declared, "no declarations parsed"
    assert declared == set(_lib.exported_symbols())
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.cqs_abi_version() == 2


# ------------------------------------------------------------------ MFMA block fusion (packing checked by the model)
def test_block_payload_round_trip():
    rng = np.random.RandomState(4)
    for k in (4, 5):
        u = np.linalg.qr(rng.randn(2**k, 2**k) + 1j * rng.randn(2**k, 2**k))[0]
        back = tape.block_matrix_from_payload(tape.block_payload(u), k)
        assert np.max(np.abs(back - u)) < 2**-21


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("L,coalesce", [(9, 4), (10, 3), (11, 5)])
def test_pass_model_with_mfma_blocks(seed, L, coalesce):
    rng = np.random.RandomState(50 + seed)
    n = 11
    c, qs = _random_circuit(n, 150, rng)
    packed = _compile_sv(c, qs, L, coalesce)
    kinds = set(int(k) for k in packed.ops_arr["kind"])
    psi0 = np.zeros(2**n, dtype=np.complex128)
    psi0[0] = 1
    got = run_pass_model(packed, psi0)
    want = O.simulate_state_vector(to_oracle_ops(c, qs), n)
    assert np.max(np.abs(got - want)) < 2e-6
    assert fidelity(got, want) > 1 - 1e-9


def test_grid_circuit_fuses_into_blocks():
    c = generators.grid_circuit(4, 5, 12, seed=1)
    qs = sorted(c.all_qubits())
    packed = _compile_sv(c, qs, 10, 4)
    kinds = [int(k) for k in packed.ops_arr["kind"]]
    assert kinds.count(5) > 0  # MFMA blocks present
    psi0 = np.zeros(2**20, dtype=np.complex128)
    psi0[0] = 1
    got = run_pass_model(packed, psi0)
    want = O.simulate_state_vector(to_oracle_ops(c, qs), 20)
    assert np.max(np.abs(got - want)) < 2e-6
