"""The rule by which a chain of passes from a basis state skips tiles (passes.hip run_passes): before
pass k the state is zero wherever the bits that no earlier pass had local differ from the basis
state's, because every op of a pass acts within its tile and maps zero to zero.  Checked on the CPU
pass model for compiled grid programs from several basis states (the GPU runs of the same programs
are compared with the oracle in tests/test_gpu_specialize.py)."""

import types

import numpy as np
import pytest

from cirq_amd import generators, tape
from tests.helpers import run_pass_model


@pytest.mark.parametrize("rows,cols,depth,L,basis", [(4, 4, 8, 10, 0), (3, 5, 10, 9, 12345), (4, 4, 12, 11, 2**15 + 77)])
def test_state_is_zero_where_the_rule_skips(rows, cols, depth, L, basis):
    circ = generators.grid_circuit(rows, cols, depth, seed=3)
    qs = sorted(circ.all_qubits())
    n = len(qs)
    low = tape.Lowering(qs, False)
    low.lower_circuit(circ)
    packed = tape.compile_ops(low.ops, n, L=L, coalesce=3, blocks=True, kmax=4, variant_bits=2)
    st = np.zeros(1 << n, dtype=np.complex128)
    st[basis] = 1
    fixed = (1 << n) - 1
    idx = np.arange(1 << n)
    skipped = 0
    for i in range(int(packed.n_passes)):
        ps = packed.pass_arr[i]
        local = 0
        for b in range(int(ps["L"])):
            local |= 1 << int(ps["lpos"][b])
        zmask = fixed & ~local
        zero = (idx & zmask) != (basis & zmask)
        assert np.all(st[zero] == 0), f"pass {i}"
        skipped += int(zero.sum())
        fixed &= ~local
        one = types.SimpleNamespace(ops_arr=packed.ops_arr, mats_arr=packed.mats_arr, aux_arr=packed.aux_arr,
                                    n_bits=packed.n_bits, pass_arr=packed.pass_arr[i : i + 1])
        st = run_pass_model(one, st)
    assert skipped > 0  # the rule skips something in these programs


def _masks(packed, n):
    """(zmask, wmask) of every pass, as run_passes sets them for a run from a basis state."""
    fixed = (1 << n) - 1
    locals_ = []
    for ps in packed.pass_arr[: int(packed.n_passes)]:
        local = 0
        for b in range(int(ps["L"])):
            local |= 1 << int(ps["lpos"][b])
        locals_.append(local)
    out = []
    for i, local in enumerate(locals_):
        zmask = fixed & ~local
        fixed &= ~local
        wmask = fixed & ~locals_[i + 1] if i + 1 < len(locals_) else 0
        out.append((zmask, wmask))
    return out


@pytest.mark.parametrize("rows,cols,depth,L,basis", [(4, 4, 8, 6, 0), (3, 5, 10, 5, 12345), (4, 4, 12, 7, 2**15 + 77)])
def test_unwritten_tiles_are_never_read(rows, cols, depth, L, basis):
    """A pass leaves unwritten the tiles the next pass neither reads (wmask): run on a buffer of NaN,
    every element a pass reads was written by the pass before it, and the final state is whole and
    equals the run that writes every tile."""
    circ = generators.grid_circuit(rows, cols, depth, seed=5)
    qs = sorted(circ.all_qubits())
    n = len(qs)
    low = tape.Lowering(qs, False)
    low.lower_circuit(circ)
    packed = tape.compile_ops(low.ops, n, L=L, coalesce=3, blocks=True, kmax=4, variant_bits=2)
    want = np.zeros(1 << n, dtype=np.complex128)
    want[basis] = 1
    buf = np.full(1 << n, np.nan, dtype=np.complex128)
    idx = np.arange(1 << n)
    unwritten = 0
    for i, (zmask, wmask) in enumerate(_masks(packed, n)):
        read = (idx & zmask) == (basis & zmask)
        if i == 0:
            st = want.copy()  # the first pass builds the basis state in registers
        else:
            assert not np.isnan(buf[read]).any(), f"pass {i} reads an element nobody wrote"
            st = np.where(read, buf, 0)
        one = types.SimpleNamespace(ops_arr=packed.ops_arr, mats_arr=packed.mats_arr, aux_arr=packed.aux_arr,
                                    n_bits=packed.n_bits, pass_arr=packed.pass_arr[i : i + 1])
        st = run_pass_model(one, st)
        want = run_pass_model(one, want)
        written = (idx & wmask) == (basis & wmask)
        buf[written] = st[written]
        unwritten += int((~written).sum())
    assert unwritten > 0  # the rule leaves something unwritten in these programs
    assert not np.isnan(buf).any()
    np.testing.assert_allclose(buf, want, atol=1e-12)
