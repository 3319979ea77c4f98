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
