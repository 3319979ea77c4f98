"""GPU test of the autotuner: it times candidate programs, keeps one no slower than the default,
and the simulator's result with the tuned program matches the oracle."""

import sys

import numpy as np
import pytest

import cirq_amd as cirq
from cirq_amd import generators
from cirq_amd.autotune import Autotuner, default_config
from oracle import cirq_np as O
from tests.helpers import fidelity, to_oracle_ops

pytestmark = pytest.mark.gpu


def test_autotuned_simulation_matches_oracle(gpu):
    print("autotune 18q grid circuit", flush=True)
    sys.stdout.flush()
    c = generators.grid_circuit(3, 6, 14, seed=3)
    qs = sorted(c.all_qubits())
    tuner = Autotuner(reps=1)
    sim = cirq.Simulator(autotune=tuner)
    psi = sim.simulate(c).final_state_vector
    want = O.simulate_state_vector(to_oracle_ops(c, qs), len(qs))
    assert np.max(np.abs(psi - want)) < 1e-5 and fidelity(psi, want) > 1 - 1e-6
    (key, log), = tuner.log.items()
    assert len(log) >= 8
    times = dict(log)
    chosen = tuner.cache[key]
    assert times[chosen] <= times[default_config(len(qs))]
    # a second simulator with the same tuner reuses the choice without timing again
    n_logged = len(tuner.log)
    psi2 = cirq.Simulator(autotune=tuner).simulate(generators.grid_circuit(3, 6, 14, seed=4)).final_state_vector
    assert len(tuner.log) == n_logged and abs(np.vdot(psi2, psi2) - 1) < 1e-5
