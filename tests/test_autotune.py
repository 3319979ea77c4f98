"""CPU tests of the autotuner's host side: every candidate configuration compiles to a program
that computes the same state (pass model vs the oracle), structure keys, and the JSON cache."""

import dataclasses

import numpy as np
import pytest

from cirq_amd import generators
from cirq_amd.autotune import Autotuner, TuneConfig, _axes, default_config, structure_key
from cirq_amd.tape import Lowering, compile_ops
from oracle import cirq_np as O
from tests.helpers import fidelity, run_pass_model, to_oracle_ops


def _lowered(rows=3, cols=4, depth=10, seed=0):
    c = generators.grid_circuit(rows, cols, depth, seed=seed)
    qs = sorted(c.all_qubits())
    low = Lowering(qs, dm=False)
    low.lower_circuit(c)
    return c, qs, low.ops


def test_every_candidate_computes_the_same_state():
    c, qs, ops = _lowered()
    n = len(qs)
    want = O.simulate_state_vector(to_oracle_ops(c, qs), n)
    base = TuneConfig(tile_bits=10, coalesce_bits=4)
    cfgs = {base}
    for name, values in _axes(n, base):
        for v in values:
            cfgs.add(dataclasses.replace(base, **{name: v}))
    assert len(cfgs) >= 10
    psi0 = np.zeros(2**n, dtype=np.complex128)
    psi0[0] = 1
    for cfg in sorted(cfgs, key=repr):
        packed = compile_ops(ops, n, **cfg.compile_kwargs(n))
        got = run_pass_model(packed, psi0)
        assert np.max(np.abs(got - want)) < 1e-5 and fidelity(got, want) > 1 - 1e-6, cfg


def test_knobs_change_the_program():
    c, qs, ops = _lowered(3, 4, 12)
    n = len(qs)
    a = compile_ops(ops, n, **TuneConfig(tile_bits=10).compile_kwargs(n))
    b = compile_ops(ops, n, **TuneConfig(tile_bits=10, kmax=4, variant_bits=0).compile_kwargs(n))
    c2 = compile_ops(ops, n, **TuneConfig(tile_bits=9).compile_kwargs(n))
    assert a.n_passes != c2.n_passes or not np.array_equal(a.ops_arr, c2.ops_arr)
    assert not np.array_equal(a.ops_arr["k"], b.ops_arr["k"]) or a.n_passes != b.n_passes


def test_structure_key():
    _, qs, ops = _lowered(seed=0)
    _, _, ops1 = _lowered(seed=1)  # other gates, same layout: the compiler sees the same shape
    _, _, ops2 = _lowered(depth=11)
    assert structure_key(ops, len(qs)) == structure_key(ops1, len(qs))
    assert structure_key(ops, len(qs)) != structure_key(ops2, len(qs))
    assert structure_key(ops, len(qs)) != structure_key(ops, len(qs) + 1)


def test_default_config_and_cache_file(tmp_path):
    assert default_config(30) == TuneConfig(tile_bits=12, coalesce_bits=4, kmax=4)
    assert default_config(32, dm=True).kmax == 5
    assert default_config(16).tile_bits == 10 and default_config(6).tile_bits == 6
    path = str(tmp_path / "tune.json")
    t = Autotuner(path)
    _, qs, ops = _lowered()
    key = structure_key(ops, len(qs))
    cfg = TuneConfig(tile_bits=11, mfma_weight=0.7)
    t.cache[key] = cfg
    t._save()
    t2 = Autotuner(path)
    assert t2.config_for(ops, len(qs)) == cfg  # served from the file, no GPU needed


def test_specialised_programs_plan_with_their_cost_entries():
    """Programs that will run as specialised kernels are planned with SPECIALIZED_COSTS (a 3-qubit
    VALU dense op is dearer than a 4-qubit MFMA block there); the interpreter keeps the table."""
    from cirq_amd.tape import COSTS, SPECIALIZED_COSTS

    cfg = TuneConfig(tile_bits=12)
    plain = cfg.compile_kwargs(20)["costs"]
    spec = cfg.compile_kwargs(20, specialized=True)["costs"]
    assert plain == COSTS
    for k, v in SPECIALIZED_COSTS.items():
        assert spec[k] == v
    assert spec["dense3"] > spec["mfma4"]
    assert {k: v for k, v in spec.items() if k not in SPECIALIZED_COSTS} == {
        k: v for k, v in plain.items() if k not in SPECIALIZED_COSTS}
