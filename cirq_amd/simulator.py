"""Drop-in Simulator / DensityMatrixSimulator running on one MI355X.

Interface and behaviour follow the reference's simulator ABCs:
  SimulatesSamples.run / run_sweep        cirq-core/cirq/sim/simulator.py:62-116
  SimulatesFinalState.simulate(_sweep)     cirq-core/cirq/sim/simulator.py:456-537
  SimulatesExpectationValues              cirq-core/cirq/sim/simulator.py:291-440,
                                          cirq-core/cirq/sim/sparse_simulator.py:193-218
  SimulatesAmplitudes                     cirq-core/cirq/sim/simulator.py:188-216
and SimulatorBase._run's split of a unitary prefix from terminal measurements
(cirq-core/cirq/sim/simulator_base.py:216-273).  The state lives on the GPU for the whole
simulation; circuits are compiled once into pass programs and re-used across resolvers.
"""

from __future__ import annotations

import collections
import math
import time
import warnings
import weakref
from typing import Any, Iterator, Sequence

import numpy as np
import torch

from . import engine
from .channels import NoiseModel
from .circuits import Circuit, Moment, expand_for_simulation
from .gates import is_parameterized
from .paulis import PauliSum, pauli_masks
from .qubits import QubitOrder
from .study import ParamResolver, ResultDict, to_resolvers
from .tape import (
    COSTS,
    DEFAULT_KMAX,
    SPECIALIZED_COSTS,
    Lowering,
    Measurement,
    ParamEvaluator,
    UnsupportedOperation,
    _has_channel,
    compile_ops,
    fill_param_mats,
    kraus_of,
)
from .trajectories import MAX_BATCH_CHANNEL_QUBITS, MAX_KRAUS, TrajectoryBatch, kraus_draws


def parse_random_state(seed) -> np.random.RandomState:
    """cirq.value.parse_random_state (cirq-core/cirq/value/random_state.py:42)."""
    if seed is None:
        return np.random.mtrand._rand
    if isinstance(seed, np.random.RandomState):
        return seed
    if isinstance(seed, (int, np.integer)):
        return np.random.RandomState(int(seed))
    raise TypeError(f"Argument must be of type cirq.value.RANDOM_STATE_OR_SEED_LIKE, got {type(seed)}.")


def _is_measurement_op(op) -> bool:
    g = getattr(op, "gate", None)
    return g is not None and type(g).__name__ == "MeasurementGate"


def _split_prefix(circuit, predicate):
    """split_into_matching_protocol_then_general (cirq-core/cirq/sim/simulator.py:952-983)."""
    blocked_q: set = set()
    blocked_k: set = set()
    prefix, suffix = Circuit(), Circuit()
    for moment in circuit:
        mp, gp = [], []
        for op in moment:
            qs = set(op.qubits)
            keys = {op.gate.key} if _is_measurement_op(op) else set()
            if predicate(op) and qs.isdisjoint(blocked_q) and keys.isdisjoint(blocked_k):
                mp.append(op)
            else:
                blocked_q |= qs
                blocked_k |= keys
                gp.append(op)
        if mp:
            prefix._moments.append(Moment(mp))
        if gp:
            suffix._moments.append(Moment(gp))
    return prefix, suffix


class _Factor:
    __slots__ = ("qubits",)

    def __init__(self, qubits):
        self.qubits = list(qubits)


def product_state_factors(circuit, qubits) -> list[list]:
    """Factors of Cirq's SimulationProductState after `circuit`, in the order Cirq samples them.

    Follows cirq-core/cirq/sim/simulation_product_state.py:83-130 (merge by Kronecker product in
    op-qubit order, identity gates skipped, whole SWAPs rename) and the initial map built over
    reversed(qubits) (cirq-core/cirq/sim/simulator_base.py:330-340); sampling walks unique
    factors in that map's order (cirq-core/cirq/sim/simulation_product_state.py:154-172).
    """
    states: dict = {}
    for q in reversed(qubits):
        states[q] = _Factor([q])
    for op in circuit.all_operations():
        g = getattr(op, "gate", None)
        name = type(g).__name__ if g is not None else ""
        if name == "IdentityGate":
            continue
        if name == "SwapPowGate" and not is_parameterized(g) and g.exponent % 2 == 1 and g.global_shift == 0:
            q0, q1 = op.qubits
            a0, a1 = states[q0], states[q1]
            if a0 is not a1:
                a1.qubits = [q0 if q == q1 else q for q in a1.qubits]
                a0.qubits = [q1 if q == q0 else q for q in a0.qubits]
                states[q0], states[q1] = a1, a0
            continue
        merged = None
        for q in op.qubits:
            if merged is None:
                merged = states[q]
            elif q not in merged.qubits:
                merged.qubits.extend(states[q].qubits)
        if merged is None:
            continue
        for q in merged.qubits:
            states[q] = merged
    out = []
    seen = set()
    for v in states.values():
        if id(v) not in seen:
            seen.add(id(v))
            out.append(list(v.qubits))
    return out


def valid_state_vector(state, n: int, atol: float = 1e-7):
    """An initial state vector checked as Cirq checks it (to_valid_state_vector /
    validate_normalized_state_vector, cirq-core/cirq/qis/states.py:766-897): a 1-D array of
    per-qubit basis values (integer or bool dtype, length n) names a basis state (its index is
    returned); otherwise the amplitudes, cast to complex64, must number 2^n and have
    sum |a|^2 close to 1 (np.isclose with atol 1e-7), else ValueError.  The pass kernels rely on
    this: split-FP16 MFMA blocks scale amplitudes by 2^15 and need |a| <= 1."""
    data = np.asarray(state)
    if data.ndim == 1 and data.dtype.kind != "c":
        if n == 2**n:
            raise ValueError("Because len(qid_shape) == product(qid_shape), it is ambiguous whether the given state "
                             "contains state vector amplitudes or per-qudit computational basis values. In this "
                             "situation you are required to pass in a state vector that is a numpy array with a "
                             "complex dtype.")
        if data.shape == (n,):
            if data.dtype.kind not in "?bBiu":
                raise ValueError(f"Expected a bool or int entry for each qudit in `state`, because len(state) == "
                                 f"len(qid_shape), but got dtype {data.dtype}.")
            idx = 0
            for i, v in enumerate(data.tolist()):
                if not 0 <= int(v) < 2:
                    raise ValueError(f"Qudit value {v} at index {i} is out of bounds for qudit dimension 2.")
                idx = 2 * idx + int(v)
            return idx
    data = data.astype(np.complex64, casting="unsafe").reshape(-1)
    if data.size != 2**n:
        raise ValueError(f"state_vector has incorrect size. Expected {2**n} but was {data.size}.")
    norm = np.sum(np.abs(data) ** 2)
    if not np.isclose(norm, 1, atol=atol):
        raise ValueError(f"State_vector is not normalized instead had norm {norm}")
    return data


def valid_density_matrix(state, n: int, atol: float = 1e-7):
    """An initial density matrix checked as Cirq checks it (to_valid_density_matrix /
    validate_density_matrix, cirq-core/cirq/qis/states.py:933-1035): a (2^n, 2^n) (or (2,)*2n)
    array must be Hermitian, of trace 1 and positive semidefinite (eigenvalues > -atol); anything
    else is taken as a state vector (or per-qubit basis values) and becomes |psi><psi|.  The
    eigenvalue check is a host eigendecomposition, skipped above 2^12 x 2^12."""
    N = 2**n
    if isinstance(state, np.ndarray):
        rho = state
        if n > 1 and rho.shape == (2,) * (2 * n):
            rho = rho.reshape(N, N)
        if rho.shape == (N, N):
            rho = rho.astype(np.complex64, casting="unsafe")
            if not np.allclose(rho, rho.conj().T, atol=atol):
                raise ValueError("The density matrix is not Hermitian.")
            trace = np.trace(rho)
            if not np.isclose(trace, 1.0, atol=atol):
                raise ValueError(f"Density matrix does not have trace 1. Instead, it has trace {trace}.")
            if N <= 4096 and not np.all(np.linalg.eigvalsh(rho) > -atol):
                raise ValueError("The density matrix is not positive semidefinite. This may happen for simulations "
                                 "using lower-precision `dtype` such as `np.complex64`. Try using `np.complex128` in "
                                 "such a case.")
            return rho
    v = valid_state_vector(state, n, atol)
    if not isinstance(v, np.ndarray):
        return v
    v = v.astype(np.complex128)
    return np.outer(v, np.conj(v))


class _Compiled:
    """A circuit lowered, scheduled and uploaded for one qubit order."""

    def __init__(self, circuit, qubits, dm: bool, device, L: int, coalesce: int, fuse: bool, blocks: bool = True,
                 tuner=None, specialized: bool = False):
        self.qubits = tuple(qubits)
        self.n = len(self.qubits)
        self.dm = dm
        self.n_bits = 2 * self.n if dm else self.n
        low = Lowering(self.qubits, dm)
        low.lower_circuit(circuit)
        self.lowering = low
        self.measurements: list[Measurement] = low.measurements
        if tuner is not None and low.ops:
            self.tuning = tuner.config_for(low.ops, self.n_bits, dm)
            kw = self.tuning.compile_kwargs(self.n_bits, specialized)
        else:
            self.tuning = None
            kw = dict(L=L, coalesce=coalesce, blocks=blocks, kmax=DEFAULT_KMAX[bool(dm)])
            if specialized:
                kw["costs"] = dict(COSTS, **SPECIALIZED_COSTS)
        self.specialize_s = None  # seconds spent compiling specialised kernels, when it does
        self.packed = compile_ops(low.ops, self.n_bits, fuse=fuse, **kw)
        self.evaluator = ParamEvaluator(low.param_refs)
        self.has_params = bool(self.packed.param_entries)
        self.dev = engine.DeviceProgram(self.packed, device)

    def mats_for(self, resolvers) -> np.ndarray:
        symvals = self.evaluator.symbol_values(resolvers)
        return fill_param_mats(self.packed, self.evaluator, symvals)


class _QubitsOnly:
    """The qubit bookkeeping of a _Compiled, for states that several programs act on."""

    def __init__(self, qubits, dm: bool = False):
        self.qubits = tuple(qubits)
        self.n = len(self.qubits)
        self.n_bits = 2 * self.n if dm else self.n


class _SimulatorBase:
    _dm = False

    def __init__(
        self,
        *,
        dtype=np.complex64,
        noise=None,
        seed=None,
        split_untangled_states: bool = True,
        tile_bits: int | None = None,
        coalesce_bits: int | None = None,
        fuse: bool = True,
        mfma_blocks: bool = True,
        autotune=False,
        trajectory_batch: int | None = None,
        specialize_passes: bool = False,
    ):
        """Cirq's arguments (dtype, noise, seed, split_untangled_states) plus engine settings:
        tile_bits / coalesce_bits / mfma_blocks fix the pass compiler's tile and fusion choices;
        autotune=True (or an autotune.Autotuner) instead times candidate settings on the GPU the
        first time each circuit structure is compiled and keeps the fastest; trajectory_batch caps
        how many repetitions of a circuit with mid-circuit measurements or state-vector noise run
        together (None: as many as a quarter of free device memory holds; 1: one at a time);
        specialize_passes=True compiles each circuit's passes to straight-line kernels the first time
        it runs (cirq_amd.specialize: seconds to a minute of hipcc, cached on disk), which pays for
        circuits simulated many times on large states."""
        if np.dtype(dtype).kind != "c":
            raise ValueError(f"dtype must be a complex type but was {dtype}")
        if np.dtype(dtype) != np.complex64:
            raise ValueError("this simulator computes in complex64; dtype must be np.complex64")
        self._dtype = np.complex64
        self.noise = NoiseModel.from_noise_model_like(noise)
        self._prng = parse_random_state(seed)
        self._split = bool(split_untangled_states)
        self._tile_bits = tile_bits
        self._coalesce = coalesce_bits
        self._fuse = fuse
        self._blocks = mfma_blocks
        if autotune is True:
            from .autotune import default_autotuner

            autotune = default_autotuner()
        self._tuner = autotune or None
        self._traj_batch = trajectory_batch
        self._specialize = bool(specialize_passes)
        self._cache: "collections.OrderedDict" = collections.OrderedDict()
        self._split_memo: dict = {}
        self._noisy_memo: dict = {}
        self._device = None
        self._sampler = None
        self._ws = None

    # ---------------------------------------------------------------- plumbing
    def _dev(self):
        if self._device is None:
            self._device = engine.require_gpu()
            self._sampler = engine.Sampler(self._device)
            self._ws = engine.Workspace(self._device)
        return self._device

    def _noisy(self, circuit) -> Circuit:
        """The circuit as simulated: sub-circuits and ops that only decompose expanded
        (circuits.expand_for_simulation), then the noise model's moments.  Memoised per circuit
        object, so repeated calls on one program reuse its compiled form."""
        hit = self._noisy_memo.get(id(circuit))
        if hit is not None and hit[0]() is circuit:
            return hit[1]
        out = expand_for_simulation(circuit)
        if not (self.noise is None or type(self.noise).__name__ == "_NoNoise"):
            qubits = sorted(out.all_qubits())
            out = Circuit.from_moments(*self.noise.noisy_moments(list(out), qubits))
        if out is not circuit:
            try:
                self._noisy_memo[id(circuit)] = (weakref.ref(circuit), out)
            except TypeError:
                return out
            while len(self._noisy_memo) > 16:
                self._noisy_memo.pop(next(iter(self._noisy_memo)))
        return out

    def _tile_params(self, n_bits: int) -> tuple[int, int]:
        if self._tile_bits is not None:
            L = self._tile_bits
        else:
            L = 12 if n_bits >= 20 else (10 if n_bits >= 14 else n_bits)
        c = self._coalesce if self._coalesce is not None else 4
        return min(L, n_bits), min(c, L, n_bits)

    def _compile(self, circuit, qubits) -> _Compiled:
        key = (id(circuit), tuple(qubits))
        hit = self._cache.get(key)
        if hit is not None and hit[0]() is circuit:
            self._cache.move_to_end(key)
            return hit[1]
        nb = 2 * len(qubits) if self._dm else len(qubits)
        engine.check_state_fits(nb, self._dev())
        L, c = self._tile_params(nb)
        comp = _Compiled(circuit, qubits, self._dm, self._dev(), L, c, self._fuse, self._blocks, self._tuner,
                         self._specialize)
        if self._specialize and comp.packed.n_passes:
            t0 = time.perf_counter()
            comp.dev.specialize()
            comp.specialize_s = time.perf_counter() - t0
        try:
            ref = weakref.ref(circuit)
        except TypeError:
            ref = lambda c=circuit: c
        self._cache[key] = (ref, comp)
        while len(self._cache) > 16:
            self._cache.popitem(last=False)
        return comp

    def _initial_index(self, initial_state, n: int) -> int:
        if initial_state is None:
            return 0
        if isinstance(initial_state, (int, np.integer)):
            v = int(initial_state)
            if v < 0 or v >= (1 << n):
                raise ValueError(f"initial_state {v} out of range for {n} qubits")
            return v
        raise TypeError("initial_state must be an int or a state array")

    def _prepare(self, comp: _Compiled, initial_state) -> torch.Tensor:
        dev = self._dev()
        n = comp.n
        if initial_state is not None and not isinstance(initial_state, (int, np.integer)):
            if self._dm:
                arr = valid_density_matrix(initial_state, n)
            else:
                arr = valid_state_vector(initial_state, n)
            if not isinstance(arr, np.ndarray):  # per-qubit basis values
                return self._prepare(comp, int(arr))
            return torch.from_numpy(np.ascontiguousarray(arr.reshape(-1), dtype=np.complex64)).to(dev)
        b = self._initial_index(initial_state, n)
        if self._dm:
            b = b * (2**n) + b
        return engine.new_state(comp.n_bits, dev, b)

    def _run_program(self, comp: _Compiled, resolvers: list, initial_state) -> Iterator[torch.Tensor]:
        mats = comp.mats_for(resolvers) if comp.has_params else None
        basis = init = None
        if initial_state is None or isinstance(initial_state, (int, np.integer)):
            basis = self._initial_index(initial_state, comp.n)
            if self._dm:
                basis = basis * (2**comp.n) + basis
        else:  # validated and uploaded once for all resolvers
            init = self._prepare(comp, initial_state)
        for i, _ in enumerate(resolvers):
            if mats is not None:
                comp.dev.set_mats(mats[i])
            if basis is None:
                st = init.clone() if i < len(resolvers) - 1 else init
                comp.dev.run(st)
            else:  # the first pass builds the basis state itself
                st = engine.empty_state(comp.n_bits, self._dev())
                comp.dev.run(st, basis=basis)
            yield st

    @staticmethod
    def _check_resolved(resolvers, comp: _Compiled):
        names = {s.name for s in comp.evaluator.symbols}
        for r in resolvers:
            for s in comp.evaluator.symbols:
                v = r.value_of(s)
                if not isinstance(v, (int, float, complex, np.number)):
                    raise ValueError(f"Circuit contains ops whose symbols were not specified in parameter sweep: {names}")

    # ---------------------------------------------------------------- sampling
    def _sample_factors(self, st, comp: _Compiled, factors, measured: list, reps: int) -> np.ndarray:
        """Samples `measured` qubits factor by factor with the simulator's prng; columns in measured order."""
        mset = set(measured)
        columns, order = [], []
        idx = {q: i for i, q in enumerate(comp.qubits)}
        for fq in factors:
            qs = [q for q in fq if q in mset]
            if not qs:
                continue
            u = self._prng.random_sample(reps)
            axes = [idx[q] for q in qs]
            if axes == list(range(comp.n)):
                ind = self._sampler.draw(st, comp.n, self._dm, u)
            elif len(axes) == comp.n:
                ind = self._sampler.draw(st, comp.n, self._dm, u, axes=axes)
            else:
                probs = self._sampler.marginal(st, comp.n, self._dm, axes)
                ind = self._sampler.draw(None, len(axes), False, u, probs=probs)
            columns.append(self._sampler.bits(ind, len(axes)))
            order += qs
        stacked = np.column_stack(columns) if columns else np.zeros((reps, 0), dtype=np.uint8)
        pos = {q: i for i, q in enumerate(order)}
        return stacked[:, [pos[q] for q in measured]]

    def _records(self, bits: np.ndarray, measured: list, meas_ops) -> dict:
        qi = {q: i for i, q in enumerate(measured)}
        results: dict = {}
        for op in meas_ops:
            g = op.gate
            out = np.zeros((bits.shape[0], len(op.qubits)), dtype=np.uint8)
            inv = g.full_invert_mask() if hasattr(g, "full_invert_mask") else ()
            for i, q in enumerate(op.qubits):
                out[:, i] = bits[:, qi[q]]
                if i < len(inv) and inv[i]:
                    out[:, i] ^= 1
            results.setdefault(g.key, []).append(out)
        return {k: np.stack(v, axis=1) for k, v in results.items()}

    # ---------------------------------------------------------------- SimulatesSamples
    def run(self, program, param_resolver=None, repetitions: int = 1) -> ResultDict:
        return self.run_sweep(program, param_resolver, repetitions)[0]

    def run_sweep(self, program, params, repetitions: int = 1) -> list[ResultDict]:
        return list(self.run_sweep_iter(program, params, repetitions))

    def _split_cached(self, circuit):
        """_split_prefix, reusing the result for a circuit object seen before (repeated run_sweep
        calls on one program then also reuse its compiled prefix)."""
        hit = self._split_memo.get(id(circuit))
        if hit is not None and hit[0]() is circuit:
            return hit[1], hit[2]
        prefix, suffix = _split_prefix(circuit, self._can_be_in_run_prefix)
        try:
            ref = weakref.ref(circuit)
        except TypeError:
            return prefix, suffix
        self._split_memo[id(circuit)] = (ref, prefix, suffix)
        while len(self._split_memo) > 16:
            self._split_memo.pop(next(iter(self._split_memo)))
        return prefix, suffix

    def sample_expectation_values(self, program, observables, *, num_samples: int, params=None,
                                  permit_terminal_measurements: bool = False) -> list[list[float]]:
        """Expectation values estimated from `num_samples` samples per observable group
        (cirq-core/cirq/work/sampler.py:311-400); see cirq_amd.observables."""
        from .observables import sample_expectation_values

        return sample_expectation_values(self, program, observables, num_samples=num_samples, params=params,
                                         permit_terminal_measurements=permit_terminal_measurements)

    def sample(self, program, *, repetitions: int = 1, params=None):
        import pandas as pd

        rows = []
        for r in self.run_sweep(program, params, repetitions):
            d = {}
            for k, v in r.measurements.items():
                w = 1 << np.arange(v.shape[1] - 1, -1, -1, dtype=np.int64)
                d[k] = (v.astype(np.int64) * w).sum(axis=1)
            for k, v in r.params.param_dict.items():
                d[str(k)] = [v] * repetitions
            rows.append(pd.DataFrame(d))
        return pd.concat(rows, ignore_index=True)

    def _can_be_in_run_prefix(self, op) -> bool:
        """Unitary (or still symbolic) ops for Simulator, any non-measurement for the DM simulator."""
        if _is_measurement_op(op):
            return False
        if self._dm or is_parameterized(op):
            return True
        g = getattr(op, "gate", None)
        obj = g if g is not None else op
        f = getattr(obj, "_has_unitary_", None)
        return bool(f()) if f is not None else False

    def run_sweep_iter(self, program, params, repetitions: int = 1) -> Iterator[ResultDict]:
        if not program.has_measurements():
            raise ValueError("Circuit has no measurements to sample.")
        if repetitions < 0:
            raise ValueError(f"Number of repetitions cannot be negative. Was {repetitions}")
        resolvers = list(to_resolvers(params))
        circuit = self._noisy(program)
        qubits = tuple(sorted(circuit.all_qubits()))
        prefix, suffix = self._split_cached(circuit)
        meas_ops = list(suffix.all_operations())
        if not all(_is_measurement_op(op) for op in meas_ops):
            yield from self._run_trajectories(prefix, suffix, qubits, resolvers, repetitions)
            return
        if repetitions == 0:
            for r in resolvers:
                yield ResultDict(params=r, records={op.gate.key: np.empty([0, 1, 1]) for op in meas_ops})
            return
        comp = self._compile(prefix, qubits)
        self._check_resolved(resolvers, comp)
        measured: list = []
        seen = set()
        for op in meas_ops:
            for q in op.qubits:
                if q not in seen:
                    seen.add(q)
                    measured.append(q)
        factors = product_state_factors(prefix, qubits) if self._split else [list(qubits)]
        for r, st in zip(resolvers, self._run_program(comp, resolvers, None)):
            bits = self._sample_factors(st, comp, factors, measured, repetitions)
            yield ResultDict(params=r, records=self._records(bits, measured, meas_ops))

    def _run_trajectories(self, prefix, suffix, qubits, resolvers, repetitions: int):
        """Per-repetition simulation of a general suffix after a deterministic prefix, as
        SimulatorBase._run does (cirq-core/cirq/sim/simulator_base.py:248-273): the prefix state is
        computed once, every repetition runs the suffix on a copy, and records of a key measured
        several times are padded to the same count.  Repetitions run in batches of copies
        (trajectories.TrajectoryBatch) unless trajectory_batch is 1 or a channel is too wide for
        the batched kernels."""
        comp = self._compile(prefix, qubits)
        self._check_resolved(resolvers, comp)
        actions = self._actions(suffix, qubits)
        keys = []
        for op in suffix.all_operations():
            if _is_measurement_op(op) and op.gate.key not in keys:
                keys.append(op.gate.key)
        batch = None
        if self._traj_batch != 1 and repetitions > 0 and TrajectoryBatch.supported(actions, self._dm):
            batch = self._trajectory_batch(qubits, actions)
        for r, st0 in zip(resolvers, self._run_program(comp, resolvers, None)):
            per_key: dict = {k: [] for k in keys}
            if batch is not None:
                for rec in batch.run(st0, r, repetitions, self._prng):
                    for k in keys:
                        per_key[k].append([bits for kk, bits in rec if kk == k])
            for i in range(repetitions if batch is None else 0):
                st = st0.clone() if i < repetitions - 1 else st0
                records: dict = {}
                for _ in self._run_actions(st, qubits, actions, r, records):
                    pass
                for k in keys:
                    per_key[k].append(records.get(k, []))
            out = {}
            for k, reps in per_key.items():
                width = max((len(x) for rec in reps for x in rec), default=1)
                largest = max((len(rec) for rec in reps), default=0)
                xs = np.zeros((repetitions, largest, width), dtype=np.uint8)
                for i, rec in enumerate(reps):
                    for j, bits in enumerate(rec):
                        xs[i, j, : len(bits)] = bits
                out[k] = xs
            yield ResultDict(params=r, records=out)

    def _trajectory_batch(self, qubits, actions) -> TrajectoryBatch:
        """The batched runner of an action list, kept (with the programs it compiled for each
        batch size) as long as the action list is cached."""
        key = ("trajectories", id(actions))
        hit = self._cache.get(key)
        if hit is not None and hit[0] is actions:
            self._cache.move_to_end(key)
            return hit[1]
        batch = TrajectoryBatch(qubits, actions, self._dm, self._dev(), self._tile_params, self._fuse, self._blocks,
                                self._ws, self._traj_batch)
        self._cache[key] = (actions, batch)
        while len(self._cache) > 16:
            self._cache.popitem(last=False)
        return batch

    # ---------------------------------------------------------------- moment steps
    def simulate_moment_steps(self, circuit, param_resolver=None, qubit_order=QubitOrder.DEFAULT,
                              initial_state=None) -> Iterator:
        """One step result per moment (cirq-core/cirq/sim/simulator.py:607-639): measurements in a
        moment collapse the state and are reported in that step."""
        r = ParamResolver(param_resolver) if not isinstance(param_resolver, ParamResolver) else param_resolver
        circuit = self._noisy(circuit)
        qubits = QubitOrder.as_qubit_order(qubit_order).order_for(circuit.all_qubits())
        actions = self._actions(circuit, qubits, per_moment=True)
        for kind, a in actions:
            if kind == "program":
                self._check_resolved([r], a)
        st = self._prepare_state(qubits, initial_state)
        if len(circuit) == 0:  # one step with the initial state (SimulatorBase._core_iterator)
            yield self._step_result(st, qubits, {})
            return
        records: dict = {}
        for _ in self._run_actions(st, qubits, actions, r, records):
            step = self._step_result(st, qubits, {k: (v[0] if len(v) == 1 else np.stack(v))
                                                  for k, v in records.items()})
            records.clear()  # the running generator keeps writing into this dict
            yield step

    def _step_result(self, st, qubits, measurements):
        cls = DensityMatrixStepResult if self._dm else StateVectorStepResult
        return cls(self, st, qubits, measurements)

    def _sample_state(self, st, qubits, measured, repetitions: int, seed) -> np.ndarray:
        prng = parse_random_state(seed) if seed is not None else self._prng
        idx = {q: i for i, q in enumerate(qubits)}
        axes = [idx[q] for q in measured]
        n = len(qubits)
        u = prng.random_sample(repetitions)
        if axes == list(range(n)):
            ind = self._sampler.draw(st, n, self._dm, u)
        elif len(axes) == n:
            ind = self._sampler.draw(st, n, self._dm, u, axes=axes)
        else:
            probs = self._sampler.marginal(st, n, self._dm, axes)
            ind = self._sampler.draw(None, len(axes), False, u, probs=probs)
        return self._sampler.bits(ind, len(axes))

    def _prepare_state(self, qubits, initial_state) -> torch.Tensor:
        return self._prepare(_QubitsOnly(qubits, self._dm), initial_state)

    # ---------------------------------------------------------------- SimulatesFinalState
    def simulate(self, program, param_resolver=None, qubit_order=QubitOrder.DEFAULT, initial_state=None):
        return self.simulate_sweep(program, param_resolver, qubit_order, initial_state)[0]

    def simulate_sweep(self, program, params, qubit_order=QubitOrder.DEFAULT, initial_state=None) -> list:
        return list(self.simulate_sweep_iter(program, params, qubit_order, initial_state))

    def _measure_terminal(self, st, comp: _Compiled, meas: list) -> dict:
        """Applies measurements at the end of a simulation (collapse + renormalise)."""
        raw = self._measure(st, comp.qubits, meas)
        return {k: (v[0] if len(v) == 1 else np.stack(v)) for k, v in raw.items()}

    def _measure(self, st, qubits, meas: list) -> dict:
        """Measures in order, collapsing the state after each (measure_state_vector,
        cirq-core/cirq/sim/state_vector.py:235-322; measure_density_matrix,
        cirq-core/cirq/sim/density_matrix_utils.py:97-182): float32 marginal in NumPy's order,
        normalised, one prng.choice, projection and renormalisation.  Returns key -> bit arrays."""
        out: dict = {}
        n = len(qubits)
        n_bits = 2 * n if self._dm else n
        idx = {q: i for i, q in enumerate(qubits)}
        for m in meas:
            axes = [idx[q] for q in m.qubits]
            probs = self._sampler.marginal(st, n, self._dm, axes)
            p = probs.cpu().numpy().astype(np.float32)
            p = np.clip(p, 0, None)
            p /= p.sum()
            r = int(self._prng.choice(len(p), p=p))
            bits = [(r >> (len(axes) - 1 - k)) & 1 for k in range(len(axes))]
            mask = value = 0
            for a, b in zip(axes, bits):
                if self._dm:
                    rb, cb = 2 * n - 1 - a, n - 1 - a
                    mask |= (1 << rb) | (1 << cb)
                    value |= (b << rb) | (b << cb)
                else:
                    bb = n - 1 - a
                    mask |= 1 << bb
                    value |= b << bb
            s = 1.0 / p[r] if self._dm else 1.0 / math.sqrt(p[r])
            engine.collapse(st, n_bits, mask, value, s)
            inv = m.invert_mask
            out.setdefault(m.key, []).append(
                np.array([b ^ (1 if k < len(inv) and inv[k] else 0) for k, b in enumerate(bits)], dtype=np.uint8)
            )
        return out

    # ---------------------------------------------------------------- trajectories
    def _stochastic(self, op) -> bool:
        """Ops that consume the prng while the state evolves: measurements, and channels on a state
        vector (sampled as in cirq-core/cirq/sim/state_vector_simulation_state.py:183-257)."""
        return _is_measurement_op(op) or (not self._dm and _has_channel(op))

    def _actions(self, circuit, qubits, per_moment: bool = False) -> list:
        """The circuit as a list of actions: ('program', _Compiled) for runs of deterministic ops,
        ('op', op) for stochastic ones in circuit order, ('step', None) after every moment when
        per_moment is set."""
        key = ("actions", id(circuit), tuple(qubits), per_moment)
        hit = self._cache.get(key)
        if hit is not None and hit[0]() is circuit:
            return hit[1]
        actions: list = []
        pending: list = []

        def flush():
            if pending:
                c = Circuit.from_moments(*pending)
                nb = 2 * len(qubits) if self._dm else len(qubits)
                L, cb = self._tile_params(nb)
                actions.append(("program", _Compiled(c, qubits, self._dm, self._dev(), L, cb, self._fuse,
                                                     self._blocks)))
                pending.clear()

        for moment in circuit:
            det = [op for op in moment if not self._stochastic(op)]
            sto = [op for op in moment if self._stochastic(op)]
            if det:
                pending.append(Moment(det))
            if sto or per_moment:
                flush()
                actions += [("op", op) for op in sto]
            if per_moment:
                actions.append(("step", None))
        flush()
        try:
            ref = weakref.ref(circuit)
        except TypeError:
            ref = lambda c=circuit: c
        self._cache[key] = (ref, actions)
        return actions

    def _apply_matrix(self, st, qubits, op_qubits, matrix):
        """st <- matrix (Cirq order over op_qubits) applied to a state vector."""
        from .tape import LOp

        idx = {q: i for i, q in enumerate(qubits)}
        n = len(qubits)
        bits = tuple(n - 1 - idx[q] for q in reversed(op_qubits))
        packed = compile_ops([LOp("dense", bits, np.asarray(matrix, dtype=np.complex128))], n,
                             L=min(n, 12), coalesce=min(4, n), fuse=False, blocks=False)
        engine.DeviceProgram(packed, st.device).run(st)

    def _apply_stochastic(self, st, qubits, op, records: dict):
        if _is_measurement_op(op):
            g = op.gate
            m = Measurement(g.key, tuple(op.qubits), g.full_invert_mask())
            for k, v in self._measure(st, qubits, [m]).items():
                records.setdefault(k, []).extend(v)
            return
        obj = getattr(op, "gate", None) or op
        mix = getattr(obj, "_mixture_", None)
        mixture = mix() if mix is not None else None
        n = len(qubits)
        narrow = len(op.qubits) <= MAX_BATCH_CHANNEL_QUBITS  # the one-copy case of the batched kernels
        idx = {q: i for i, q in enumerate(qubits)}
        targets = [n - 1 - idx[q] for q in reversed(op.qubits)]
        if mixture is not None and mixture is not NotImplemented:
            probabilities, unitaries = zip(*mixture)
            index = self._prng.choice(range(len(unitaries)), p=probabilities)
            if narrow:
                engine.apply_select_batch(st, n, 1, targets, np.stack([np.asarray(u) for u in unitaries]),
                                          np.array([index]), np.ones(1, dtype=np.float32))
            else:
                self._apply_matrix(st, qubits, op.qubits, unitaries[index])
            return
        kraus = kraus_of(obj)
        p = self._prng.random()
        if narrow and len(kraus) <= MAX_KRAUS:
            mats = np.stack([np.asarray(K) for K in kraus])
            W = engine.kraus_weights_batch(st, n, 1, targets, mats, self._ws)
            sel, w = kraus_draws(W, np.array([p]))
            engine.apply_select_batch(st, n, 1, targets, mats, sel, np.array([1.0 / math.sqrt(w[0])], dtype=np.float32))
            return
        fallback_weight, fallback_index = 0.0, 0
        weight, buf = None, None
        for index, K in enumerate(kraus):
            buf = st.clone()
            self._apply_matrix(buf, qubits, op.qubits, K)
            weight = engine.norm2(buf, n, self._ws)
            if weight > fallback_weight:
                fallback_index, fallback_weight = index, weight
            p -= weight
            if p < 0:
                break
        if p >= 0 or weight == 0:
            buf = st.clone()
            self._apply_matrix(buf, qubits, op.qubits, kraus[fallback_index])
            weight = fallback_weight
        engine.scale(buf, n, 1.0 / math.sqrt(weight))
        st.copy_(buf)

    def _run_actions(self, st, qubits, actions, resolver, records: dict):
        """Runs actions on st; yields at every ('step', None)."""
        for kind, a in actions:
            if kind == "program":
                if a.has_params:
                    a.dev.set_mats(a.mats_for([resolver])[0])
                a.dev.run(st)
            elif kind == "op":
                self._apply_stochastic(st, qubits, a, records)
            else:
                yield

    def _needs_trajectory(self, circuit) -> bool:
        if circuit.has_measurements() and not circuit.are_all_measurements_terminal():
            return True
        return not self._dm and any(_has_channel(op) for op in circuit.all_operations())

    def simulate_sweep_iter(self, program, params, qubit_order=QubitOrder.DEFAULT, initial_state=None):
        resolvers = list(to_resolvers(params))
        circuit = self._noisy(program)
        qubits = QubitOrder.as_qubit_order(qubit_order).order_for(circuit.all_qubits())
        if self._needs_trajectory(circuit):
            actions = self._actions(circuit, qubits)
            for r in resolvers:
                st = self._prepare_state(qubits, initial_state)
                records: dict = {}
                for _ in self._run_actions(st, qubits, actions, r, records):
                    pass
                meas = {k: (v[0] if len(v) == 1 else np.stack(v)) for k, v in records.items()}
                yield self._trial_result(r, meas, st, _QubitsOnly(qubits))
            return
        body, meas = _split_terminal_measurements(circuit)
        comp = self._compile(body, qubits)
        self._check_resolved(resolvers, comp)
        for r, st in zip(resolvers, self._run_program(comp, resolvers, initial_state)):
            measurements = self._measure_terminal(st, comp, meas) if meas else {}
            yield self._trial_result(r, measurements, st, comp)

    # ---------------------------------------------------------------- expectation values
    def simulate_expectation_values(
        self,
        program,
        observables,
        param_resolver=None,
        qubit_order=QubitOrder.DEFAULT,
        initial_state=None,
        permit_terminal_measurements: bool = False,
    ) -> list:
        return self.simulate_expectation_values_sweep(
            program, observables, param_resolver, qubit_order, initial_state, permit_terminal_measurements
        )[0]

    def simulate_expectation_values_sweep(
        self,
        program,
        observables,
        params,
        qubit_order=QubitOrder.DEFAULT,
        initial_state=None,
        permit_terminal_measurements: bool = False,
    ) -> list[list]:
        return list(
            self.simulate_expectation_values_sweep_iter(
                program, observables, params, qubit_order, initial_state, permit_terminal_measurements
            )
        )

    def simulate_expectation_values_sweep_iter(
        self,
        program,
        observables,
        params,
        qubit_order=QubitOrder.DEFAULT,
        initial_state=None,
        permit_terminal_measurements: bool = False,
    ):
        """Expectation values of the final state of every resolver's simulation, as
        cirq-core/cirq/sim/sparse_simulator.py:193-218 computes them from simulate_sweep_iter:
        ValueError if any measurement is terminal (unless permitted); circuits with mid-circuit
        measurements or state-vector noise take one sampled trajectory per resolver."""
        if not permit_terminal_measurements and program.are_any_measurements_terminal():
            raise ValueError(
                "Provided circuit has terminal measurements, which may "
                "skew expectation values. If this is intentional, set "
                "permit_terminal_measurements=True."
            )
        if not isinstance(observables, list):
            observables = [observables]
        pslist = [PauliSum.wrap(o) for o in observables]
        resolvers = list(to_resolvers(params))
        circuit = self._noisy(program)
        qubits = QubitOrder.as_qubit_order(qubit_order).order_for(circuit.all_qubits())
        qmap = {q: i for i, q in enumerate(qubits)}
        n = len(qubits)
        terms, owner = [], []
        for oi, ps in enumerate(pslist):
            for t in ps:
                for q in t.qubits:
                    if q not in qmap:
                        raise KeyError(q)
                terms.append(pauli_masks(t, qmap, n))
                owner.append(oi)
        owner = np.array(owner, dtype=np.int64)

        def values(st):
            vals = engine.expect_terms(st, n, self._dm, terms, self._ws).cpu().numpy()
            sums = np.zeros(len(pslist), dtype=np.complex128)
            np.add.at(sums, owner, vals)
            return [self._ev_value(v) for v in sums]

        if self._needs_trajectory(circuit):
            for res in self.simulate_sweep_iter(program, resolvers, qubits, initial_state):
                yield values(res._state)
            return
        body, meas = _split_terminal_measurements(circuit)
        comp = self._compile(body, qubits)
        self._check_resolved(resolvers, comp)
        for r, st in zip(resolvers, self._run_program(comp, resolvers, initial_state)):
            if meas:
                self._measure_terminal(st, comp, meas)
            yield values(st)

    def _ev_value(self, v):
        return v

    # ---------------------------------------------------------------- amplitudes
    def compute_amplitudes(self, program, bitstrings, param_resolver=None, qubit_order=QubitOrder.DEFAULT):
        return self.compute_amplitudes_sweep(program, bitstrings, param_resolver, qubit_order)[0]

    def compute_amplitudes_sweep(self, program, bitstrings, params, qubit_order=QubitOrder.DEFAULT) -> list:
        return list(self.compute_amplitudes_sweep_iter(program, bitstrings, params, qubit_order))

    def compute_amplitudes_sweep_iter(self, program, bitstrings, params, qubit_order=QubitOrder.DEFAULT):
        """Amplitudes of the final state at the given big-endian bitstrings, per resolver
        (cirq-core/cirq/sim/state_vector_simulator.py:74-98): read on the GPU from the final
        state, with the same renormalisation final_state_vector applies."""
        if self._dm:
            raise NotImplementedError("amplitudes are defined for state vectors only")
        if isinstance(bitstrings, np.ndarray) and len(bitstrings.shape) > 1:
            raise ValueError("The list of bitstrings must be input as a 1-dimensional array of ints. Got an array "
                             f"with shape {bitstrings.shape}.")
        idx = torch.tensor([int(b) for b in bitstrings], dtype=torch.int64)
        for res in self.simulate_sweep_iter(program, params, qubit_order):
            amps = res._state[idx.to(res._state.device)].cpu().numpy()
            yield (amps * np.float32(res._renorm_factor())).tolist()


def _split_terminal_measurements(circuit):
    """(circuit without terminal measurements, terminal measurements in order)."""
    meas: list[Measurement] = []
    if not circuit.has_measurements():
        return circuit, meas
    if not circuit.are_all_measurements_terminal():
        raise NotImplementedError("mid-circuit measurements are not supported by this simulator")
    body = Circuit()
    for m in circuit:
        ops = []
        for op in m:
            if _is_measurement_op(op):
                g = op.gate
                meas.append(Measurement(g.key, tuple(op.qubits), g.full_invert_mask()))
            else:
                ops.append(op)
        body._moments.append(Moment(ops))
    return body, meas


# --------------------------------------------------------------------------------------------
class StateVectorTrialResult:
    """Final state of a simulation (cirq-core/cirq/sim/state_vector_simulator.py:106-146)."""

    def __init__(self, params, measurements, state: torch.Tensor, qubits, ws):
        self.params = params
        self.measurements = measurements
        self._state = state
        self._qubits = tuple(qubits)
        self.qubit_map = {q: i for i, q in enumerate(self._qubits)}
        self._ws = ws
        self._fsv = None
        self._scale = None
        self._scaled = None

    def _renorm_factor(self) -> float:
        """1/norm when rescaling brings the total probability closer to 1, else 1
        (cirq-core/cirq/sim/state_vector_simulator.py:131-146); computed once, in FP64 on the GPU."""
        if self._scale is None:
            n = len(self._qubits)
            n2 = engine.norm2(self._state, n, self._ws)
            norm = math.sqrt(n2)
            eps = math.sqrt(np.finfo(np.complex64).eps)
            self._scale = 1.0
            if abs(norm - 1) > eps:
                warnings.warn(f"final state vector's norm={norm} is too far from 1, {abs(norm - 1)} > {eps}.")
            else:
                t = self._state.clone()
                engine.scale(t, n, 1.0 / norm)
                n2b = engine.norm2(t, n, self._ws)
                if abs(n2b - 1) < abs(n2 - 1):
                    self._scale = 1.0 / norm
                    self._scaled = t
        return self._scale

    @property
    def final_state_vector(self) -> np.ndarray:
        """Copied to the host; renormalised only when that improves the total probability."""
        if self._fsv is None:
            ret = self._scaled if self._renorm_factor() != 1.0 else self._state
            self._fsv = ret.cpu().numpy()
            self._scaled = None
        return self._fsv

    def state_vector(self, copy: bool = False) -> np.ndarray:
        v = self.final_state_vector
        return v.copy() if copy else v

    def dirac_notation(self, decimals: int = 2) -> str:
        v = self.final_state_vector
        n = len(self._qubits)
        parts = []
        for i in np.flatnonzero(np.abs(v) > 10 ** (-decimals)):
            parts.append(f"({np.round(v[i], decimals)})|{i:0{n}b}⟩")
        return " + ".join(parts)

    def __repr__(self):
        return f"StateVectorTrialResult(params={self.params!r}, measurements={self.measurements!r})"


class StateVectorStepResult:
    """The state after one moment (SparseSimulatorStep, cirq-core/cirq/sim/sparse_simulator.py:221-283)."""

    def __init__(self, sim, state: torch.Tensor, qubits, measurements: dict):
        self._sim = sim
        self._state = state.clone()
        self._qubits = tuple(qubits)
        self.qubit_map = {q: i for i, q in enumerate(self._qubits)}
        self.measurements = measurements

    def state_vector(self, copy: bool = True) -> np.ndarray:
        return self._state.cpu().numpy()

    def dirac_notation(self, decimals: int = 2) -> str:
        v = self.state_vector()
        n = len(self._qubits)
        return " + ".join(f"({np.round(v[i], decimals)})|{i:0{n}b}⟩" for i in np.flatnonzero(np.abs(v) > 10 ** -decimals))

    def sample(self, qubits, repetitions: int = 1, seed=None) -> np.ndarray:
        """uint8[repetitions, len(qubits)] (sample_state_vector, cirq-core/cirq/sim/state_vector.py:170-232)."""
        return self._sim._sample_state(self._state, self._qubits, list(qubits), repetitions, seed)


class DensityMatrixStepResult(StateVectorStepResult):
    """The density matrix after one moment (cirq-core/cirq/sim/density_matrix_simulator.py:238-330)."""

    def density_matrix(self, copy: bool = True) -> np.ndarray:
        n = len(self._qubits)
        return self._state.cpu().numpy().reshape(2**n, 2**n)

    def state_vector(self, copy: bool = True):
        raise NotImplementedError("a density-matrix step has no state vector")


class DensityMatrixTrialResult:
    """Final density matrix of a simulation (cirq-core/cirq/sim/density_matrix_simulator.py:368-374)."""

    def __init__(self, params, measurements, state: torch.Tensor, qubits, ws):
        self.params = params
        self.measurements = measurements
        self._state = state
        self._qubits = tuple(qubits)
        self.qubit_map = {q: i for i, q in enumerate(self._qubits)}
        self._fdm = None

    @property
    def final_density_matrix(self) -> np.ndarray:
        if self._fdm is None:
            n = len(self._qubits)
            self._fdm = self._state.cpu().numpy().reshape(2**n, 2**n)
        return self._fdm

    def __repr__(self):
        return f"DensityMatrixTrialResult(params={self.params!r}, measurements={self.measurements!r})"


class Simulator(_SimulatorBase):
    """State-vector simulator (cirq.Simulator drop-in, cirq-core/cirq/sim/sparse_simulator.py:31-218)."""

    _dm = False

    def _trial_result(self, params, measurements, st, comp):
        return StateVectorTrialResult(params, measurements, st, comp.qubits, self._ws)


class DensityMatrixSimulator(_SimulatorBase):
    """Density-matrix simulator (cirq.DensityMatrixSimulator drop-in, cirq-core/cirq/sim/density_matrix_simulator.py:32-235)."""

    _dm = True

    def _trial_result(self, params, measurements, st, comp):
        return DensityMatrixTrialResult(params, measurements, st, comp.qubits, self._ws)

    def _ev_value(self, v):
        # Tr(rho P) is returned as a float (cirq-core/cirq/ops/pauli_string.py:764)
        return float(np.real(v))


# ---- convenience functions (cirq-core/cirq/sim/mux.py) -----------------------------------------
def final_state_vector(program, *, initial_state=0, param_resolver=None, qubit_order=QubitOrder.DEFAULT,
                       ignore_terminal_measurements: bool = False, dtype=np.complex64, seed=None) -> np.ndarray:
    if not ignore_terminal_measurements and program.has_measurements():
        raise ValueError("Circuit contains measurements; set ignore_terminal_measurements=True")
    if ignore_terminal_measurements:
        program, _ = _split_terminal_measurements(program)
    return Simulator(dtype=dtype, seed=seed).simulate(program, param_resolver, qubit_order, initial_state).final_state_vector


def final_density_matrix(program, *, noise=None, initial_state=0, param_resolver=None,
                         qubit_order=QubitOrder.DEFAULT, dtype=np.complex64, seed=None,
                         ignore_measurement_results: bool = True) -> np.ndarray:
    return DensityMatrixSimulator(dtype=dtype, noise=noise, seed=seed).simulate(
        program, param_resolver, qubit_order, initial_state
    ).final_density_matrix


def sample(program, *, noise=None, param_resolver=None, repetitions: int = 1, dtype=np.complex64, seed=None):
    sim = (DensityMatrixSimulator if noise is not None else Simulator)(dtype=dtype, noise=noise, seed=seed)
    return sim.run(program, param_resolver, repetitions)
