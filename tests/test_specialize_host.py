"""Host side of the pass specialiser (cirq_amd.specialize): the generated source mirrors the op
interpreter's dispatch for every op kind, and a small program compiles and exposes one launch
handle per pass (no GPU needed; the GPU parity is tests/test_gpu_specialize.py)."""

import numpy as np

from cirq_amd import _lib, generators, specialize, tape
from tests.test_host import _random_circuit


def _program(seed=1, n=12, L=10, blocks=True):
    rng = np.random.RandomState(seed)
    c, qs = _random_circuit(n, 160, rng)
    low = tape.Lowering(qs, False)
    low.lower_circuit(c)
    return tape.compile_ops(low.ops, n, L=L, coalesce=3, blocks=blocks, kmax=4)


def test_sources_follow_the_interpreter_dispatch():
    packed = _program()
    srcs = specialize.program_sources(packed)
    text = "".join(srcs)
    assert len(srcs) == -(-int(packed.n_passes) // specialize.KERNELS_PER_FILE)
    assert len(specialize.program_sources(packed, adjoint=False)) == len(srcs)
    for i in range(int(packed.n_passes)):
        assert f"void k_spec_{i}(" in text
    kinds = packed.ops_arr["kind"] & 0xFF
    nt, per = specialize._shape(10)
    reg = sum(specialize._is_register_dense(op, nt, per) for op in packed.ops_arr[: packed.n_ops])
    mixing = sum(specialize._mixing(op, nt, per) for op in packed.ops_arr[: packed.n_ops])
    assert text.count("reg_dense_k<") == reg
    assert text.count("tile_block_mfma<") + text.count("tile_dense<") == mixing
    assert text.count("ew_diag<") == int((kinds[: packed.n_ops] == _lib.OP_DIAG).sum())
    assert text.count("ew_pairs<") == int((kinds[: packed.n_ops] == _lib.OP_PAIRPHASE).sum())
    assert text.count("regs_to_tile_permuted<") == int((kinds[: packed.n_ops] == _lib.OP_PERMUTE).sum())


def test_outer_controls_and_variants_become_constants():
    c = generators.grid_circuit(4, 4, 16, seed=2)
    low = tape.Lowering(sorted(c.all_qubits()), False)
    low.lower_circuit(c)
    packed = tape.compile_ops(low.ops, 16, L=10, coalesce=4, blocks=True, kmax=4, variant_bits=3)
    ops = packed.ops_arr[: packed.n_ops]
    has_var = [op for op in ops if (int(op["t"][5]) >> 24) & 3]
    assert has_var  # the grid circuit's fused blocks depend on outer bits
    op = has_var[0]
    expr = specialize._variant_expr(op)
    code = int(op["t"][5])
    for i in range((code >> 24) & 3):
        assert f"(base >> {(code >> (7 * i)) & 127})" in expr
    # an op controlled by outer bits: the tile's outer bits are tested against constant masks
    rec = np.zeros(1, dtype=_lib.OP_DTYPE)[0]
    rec["co_mask_lo"], rec["co_mask_hi"], rec["co_val_lo"], rec["co_val_hi"] = 0x30, 0x1, 0x10, 0x1
    assert specialize._outer_guard(rec) == "if ((base & 0x100000030ull) == 0x100000010ull) "
    rec["co_mask_lo"] = rec["co_mask_hi"] = 0
    assert specialize._outer_guard(rec) == ""


def test_small_program_compiles_to_one_handle_per_pass(tmp_path):
    c = generators.grid_circuit(2, 3, 4, seed=0)
    low = tape.Lowering(sorted(c.all_qubits()), False)
    low.lower_circuit(c)
    packed = tape.compile_ops(low.ops, 6, L=4, coalesce=2, blocks=False)
    sp = specialize.specialize(packed, cache_dir=tmp_path)
    assert sp.path.exists() and sp.n_passes == int(packed.n_passes)
    assert all(h for h in list(sp.handles)[: sp.n_passes])
    again = specialize.specialize(packed, cache_dir=tmp_path)
    assert again is sp  # the same source is compiled once per process


def test_reverse_sweep_sources_write_one_partial_per_gradient_slot():
    from cirq_amd.adjoint import AdjointProgram

    circuit, syms, vals, ham, edges = generators.qaoa_maxcut(10, 2, seed=0)
    prog = AdjointProgram(circuit, ham, tile_bits=8, adjoint_tile_bits=8)
    text = "".join(specialize.program_sources(prog.rev, adjoint=True))
    for i in range(int(prog.rev.n_passes)):
        assert f"void k_aspec_{i}(" in text
    assert text.count("write_wave_partial<") == prog.n_slots
    kinds = prog.rev.ops_arr["kind"][: prog.rev.n_ops] & 0xFF
    grads = int(((kinds == _lib.OP_GRAD_DIAG) | (kinds == _lib.OP_GRAD_DENSE)).sum())
    assert grads == prog.n_slots
