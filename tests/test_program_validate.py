"""cqs_program_validate (host code of the C ABI, no GPU needed): every program the compiler emits
passes, and each kind of malformed program a C caller could hand over is refused with
ValueError before any kernel would read or write out of bounds."""

import numpy as np
import pytest
import sympy

import cirq_amd as cirq
from cirq_amd import _lib, engine, generators, tape
from cirq_amd.adjoint import AdjointProgram
from cirq_amd.dm_adjoint import DMAdjointProgram
from tests.test_host import _random_circuit


def _compiled(n=10, L=8, co=3, seed=0, dm=False):
    c, qs = _random_circuit(5 if dm else n, 80, np.random.RandomState(seed), with_channels=dm)
    low = tape.Lowering(qs, dm)
    low.lower_circuit(c)
    nb = 2 * len(qs) if dm else len(qs)
    return tape.compile_ops(low.ops, nb, L=min(L, nb), coalesce=min(co, L, nb))


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("tiles", [(4, 2), (8, 3), (10, 4), (12, 4), (13, 4)])
def test_compiled_programs_validate(seed, tiles):
    n = max(14, tiles[0] + 1)
    engine.validate_program(_compiled(n, *tiles, seed=seed))


@pytest.mark.parametrize("seed", range(3))
def test_density_matrix_and_grid_programs_validate(seed):
    engine.validate_program(_compiled(seed=seed, dm=True, L=6, co=2))
    c = generators.grid_circuit(5, 6, 12, seed=seed, n_qubits=22)
    low = tape.Lowering(sorted(c.all_qubits()), False)
    low.lower_circuit(c)
    for L in (12, 13):
        engine.validate_program(tape.compile_ops(low.ops, 22, L=L, coalesce=4, kmax=4))
        engine.validate_program(tape.compile_ops(low.ops, 22, L=L, coalesce=4, kmax=5))


def test_adjoint_programs_validate():
    circuit, syms, vals = generators.hardware_efficient_ansatz(12, 3, seed=1)
    psum = generators.random_pauli_sum(sorted(circuit.all_qubits()), 8, 3, seed=1)
    prog = AdjointProgram(circuit, psum)
    engine.validate_program(prog.fwd)
    engine.validate_program(prog.rev, adjoint=True)
    with pytest.raises(ValueError, match="adjoint"):
        engine.validate_program(prog.rev, adjoint=False)  # GRAD ops outside an adjoint program
    nc, syms, vals = generators.noisy_line_ansatz(4, 2, seed=0, p=0.05, gamma=0.05)
    dprog = DMAdjointProgram(nc, generators.sum_z(sorted(nc.all_qubits())))
    engine.validate_program(dprog.fwd)
    engine.validate_program(dprog.rev, adjoint=True)


def _mutate(packed, fn):
    bad = tape.PackedProgram.__new__(tape.PackedProgram)
    bad.__dict__.update(packed.__dict__)
    bad.ops_arr = packed.ops_arr.copy()
    bad.pass_arr = packed.pass_arr.copy()
    bad.aux_arr = packed.aux_arr.copy()
    fn(bad)
    return bad


def _first(packed, kind):
    idx = np.flatnonzero((packed.ops_arr["kind"] & 0xFF) == kind)
    assert len(idx), f"no op of kind {kind}"
    return int(idx[0])


def test_malformed_programs_are_refused():
    p = _compiled(14, 10, 4, seed=2)
    P = p.pass_arr.dtype.names
    dense = _first(p, _lib.OP_DENSE)
    cases = {
        "target outside the tile": lambda b: b.ops_arr["t"].__setitem__((dense, 0), 40),
        "repeated target": lambda b: b.ops_arr["t"].__setitem__((dense, slice(0, 2)), 0) if b.ops_arr["k"][dense] >= 2
        else b.ops_arr["k"].__setitem__(dense, 9),
        "matrix past the buffer": lambda b: b.ops_arr["mat"].__setitem__(dense, int(b.mats_arr.size)),
        "unknown kind": lambda b: b.ops_arr["kind"].__setitem__(dense, 77),
        "ops past the table": lambda b: b.pass_arr["op_count"].__setitem__(0, len(b.ops_arr) + 5),
        "repeated local position": lambda b: b.pass_arr["lpos"].__setitem__((0, 1), int(b.pass_arr["lpos"][0, 0])),
        "dense k above max_k": lambda b: b.pass_arr["max_k"].__setitem__(slice(None), 0),
        "outer control on a tile bit": lambda b: b.ops_arr["co_mask_lo"].__setitem__(
            dense, 1 << int(b.pass_arr["lpos"][0, 0])),
    }
    assert "lpos" in P
    for name, fn in cases.items():
        with pytest.raises(ValueError):
            engine.validate_program(_mutate(p, fn))
    # a BLOCK whose operands would start misaligned, a variant bit inside the tile
    g = generators.grid_circuit(4, 5, 10, seed=1)
    low = tape.Lowering(sorted(g.all_qubits()), False)
    low.lower_circuit(g)
    q = tape.compile_ops(low.ops, 20, L=12, coalesce=4, kmax=4)
    blk = _first(q, _lib.OP_BLOCK)
    with pytest.raises(ValueError, match="aligned"):
        engine.validate_program(_mutate(q, lambda b: b.ops_arr["mat"].__setitem__(blk, int(b.ops_arr["mat"][blk]) + 1)))
    with pytest.raises(ValueError, match="variant"):
        def inner_variant(b):
            pi = int(np.flatnonzero(b.pass_arr["op_begin"] <= blk)[-1])
            b.ops_arr["t"][blk, 5] = (1 << 24) | int(b.pass_arr["lpos"][pi, 0])
        engine.validate_program(_mutate(q, inner_variant))
