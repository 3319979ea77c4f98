"""CPU tests of the batched-sweep planner (cirq_amd.batch.BatchPrograms): the descriptor tables
and launch lists, executed launch by launch by the NumPy kernel models, must give the oracle's
expectation values and finite-difference gradients for every point of every circuit."""

import numpy as np
import pytest

import cirq_amd as cirq
from cirq_amd import _lib, generators
from cirq_amd.batch import BatchPrograms
from tests.helpers import run_batch_model
from tests.test_adjoint_host import _oracle_ev, _oracle_grad


def _workload(seed=0):
    rng = np.random.RandomState(seed)
    circuits, observables, values = [], [], []
    for k, (n, layers) in enumerate([(2, 2), (6, 1), (5, 3)]):
        c, syms, _ = generators.hardware_efficient_ansatz(n, layers, seed=k, prefix=f"c{k}_")
        qs = sorted(c.all_qubits())
        terms = list(generators.random_pauli_sum(qs, 3, 2, seed=k, include_z=False))
        circuits.append(c)
        observables.append(terms)
        values.append(rng.uniform(-np.pi, np.pi, (2 + k % 2, len(syms))))
    return circuits, observables, values


@pytest.mark.parametrize("tiles", [(4, 3), (3, 2)])
def test_batch_plan_matches_oracle(tiles):
    circuits, observables, values = _workload()
    h = BatchPrograms(circuits, observables, tile_bits=tiles[0], adjoint_tile_bits=tiles[1], coalesce_bits=1)
    h.plan(tuple(v.shape[0] for v in values))
    kinds = {k for k, *_ in h.launches}
    assert kinds == {"init", "fwd", "adj"}
    assert len({L for k, L, *_ in h.launches if k == "fwd"}) >= 2  # mixed tile sizes share levels
    mats = np.zeros(h.mats_elems, dtype=np.complex64)
    symvals = [v.T.copy() for v in values]
    h.fill_mats(symvals, mats)
    raw, slots = run_batch_model(h, mats)
    evs, grads = h.results(raw, slots, symvals)
    for c, circuit in enumerate(circuits):
        qubits = sorted(circuit.all_qubits())
        syms = h.symbols[c]
        for b in range(values[c].shape[0]):
            res = cirq.ParamResolver(dict(zip(syms, values[c][b])))
            for t, term in enumerate(observables[c]):
                want = _oracle_ev(circuit, qubits, cirq.PauliSum.wrap(term), res)
                assert abs(evs[c][b, t] - want) < 1e-5
            psum = cirq.PauliSum.from_pauli_strings(observables[c])
            np.testing.assert_allclose(grads[c][b], _oracle_grad(circuit, qubits, psum, syms, values[c][b]),
                                       atol=2e-5)


def test_batch_validate_rejects_bad_descriptors():
    circuits, observables, values = _workload()
    h = BatchPrograms(circuits, observables, tile_bits=4, adjoint_tile_bits=3, coalesce_bits=1)
    h.plan(tuple(v.shape[0] for v in values))
    lib = _lib.lib()
    kind, L, off, n, tiles = next(x for x in h.launches if x[0] == "fwd")
    good = h.items[off : off + n].copy()
    args = lambda arr, arena=h.arena_elems: (h.passes.ctypes.data, len(h.passes), len(h.ops), arr.ctypes.data,
                                             len(arr), L, arena, h.mats_elems, 0)
    assert lib.cqs_batch_validate(*args(good)) == 0
    bad = good.copy()
    bad[0]["state"] = h.arena_elems  # state outside the arena
    assert lib.cqs_batch_validate(*args(bad)) == _lib.CQS_ERR_ARG
    bad = good.copy()
    bad[-1]["tile_begin"] += 1  # gap in the workgroup numbering
    assert lib.cqs_batch_validate(*args(bad)) == _lib.CQS_ERR_ARG
    bad = good.copy()
    bad[0]["pass"] = len(h.passes)  # pass index out of range
    assert lib.cqs_batch_validate(*args(bad)) == _lib.CQS_ERR_ARG
    assert lib.cqs_batch_validate(*args(good, arena=1)) == _lib.CQS_ERR_ARG
