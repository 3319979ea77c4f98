"""Expectation values estimated from samples: Cirq's `Sampler.sample_expectation_values` and the
observable-measurement framework under it (cirq-core/cirq/work/sampler.py:311-400,
cirq-core/cirq/work/observable_measurement.py, observable_grouping.py, observable_settings.py,
observable_measurement_data.py), restated so that the same calls run on this simulator.

Observables are grouped greedily into qubit-wise commuting sets (one basis per qubit); each group
is measured by appending parameterised X / Y basis-change layers and one measurement of every
qubit (key 'z') to the circuit, and `run_sweep` is called in chunks of repetitions over all
(group, circuit parameter) pairs in the order Cirq uses.  The simulator draws every chunk's
samples with its own prng exactly as its `run_sweep` does, so given the same seed the bitstrings
follow Cirq's draw order.

Engine side: the basis-change layers are parameterised, so every group and every chunk reuses one
compiled program (only the matrix buffer changes), and each simulated state feeds a whole chunk of
draws from the GPU sampler.
"""

from __future__ import annotations

import dataclasses
import itertools
import warnings
from collections.abc import Callable, Iterable, Mapping, Sequence

import numpy as np
import sympy

from .circuits import Circuit, Moment
from .gates import XPowGate, YPowGate, measure
from .paulis import X, Y, Z, PauliString, PauliSum
from .study import ParamResolver, to_resolvers

MAX_REPETITIONS_PER_JOB = 3_000_000  # cirq-core/cirq/work/observable_measurement.py:45

# named one-qubit states of a product initial state (cirq.KET_ZERO etc.)
KET_ZERO, KET_ONE, KET_PLUS, KET_MINUS, KET_IMAG, KET_MINUS_IMAG = "0", "1", "+", "-", "i", "-i"

# (Pauli index, flipped) -> (Xf, Yf) exponents that rotate the Pauli's eigenbasis onto Z
# (cirq-core/cirq/work/observable_measurement.py:145-152)
_OBS_TO_PARAM_VAL = {
    (0, False): (0, -1 / 2),
    (0, True): (0, +1 / 2),
    (1, False): (1 / 2, 0),
    (1, True): (-1 / 2, 0),
    (2, False): (0, 0),
    (2, True): (1, 0),
}
# initial named state -> (Xi, Yi) exponents that prepare it from |0> (observable_measurement.py:158-165)
_STATE_TO_PARAM_VAL = {
    KET_PLUS: (0, +1 / 2),
    KET_MINUS: (0, -1 / 2),
    KET_IMAG: (-1 / 2, 0),
    KET_MINUS_IMAG: (+1 / 2, 0),
    KET_ZERO: (0, 0),
    KET_ONE: (1, 0),
}


@dataclasses.dataclass(frozen=True)
class InitObsSetting:
    """An initial product state (qubit -> named state) and an observable (PauliString)
    (cirq-core/cirq/work/observable_settings.py:31-60)."""

    init_state: tuple  # ((qubit, named state), ...) sorted by qubit
    observable: PauliString

    def __post_init__(self):
        init_qs = {q for q, _ in self.init_state}
        if not set(self.observable.qubits) <= init_qs:
            raise ValueError(
                "`observable`'s qubits should be a subset of those found in `init_state`. "
                f"observable qubits: {self.observable.qubits}. init_state qubits: {sorted(init_qs)}"
            )

    @property
    def init_map(self) -> dict:
        return dict(self.init_state)


def _state(mapping: Mapping) -> tuple:
    return tuple(sorted(mapping.items(), key=lambda kv: kv[0]))


def zeros_state(qubits: Iterable) -> tuple:
    return _state(dict.fromkeys(qubits, KET_ZERO))


def observables_to_settings(observables: Iterable[PauliString], qubits: Iterable):
    """Each observable measured from the all-zeros state (observable_settings.py:119-126)."""
    qubits = list(qubits)
    for obs in observables:
        yield InitObsSetting(init_state=zeros_state(qubits), observable=obs)


def _max_weight_observable(observables: Iterable[PauliString]) -> PauliString | None:
    """The Pauli at every qubit any observable acts on, or None when two disagree
    (observable_settings.py:63-87)."""
    m: dict = {}
    for obs in observables:
        for q, p in obs.items():
            if q in m:
                if m[q] != p:
                    return None
            else:
                m[q] = p
    return PauliString(qubit_pauli_map=m)


def _max_weight_state(states: Iterable[tuple]) -> tuple | None:
    """The named state at every qubit, or None when two disagree (observable_settings.py:90-112)."""
    m: dict = {}
    for st in states:
        for q, s in st:
            if q in m:
                if m[q] != s:
                    return None
            else:
                m[q] = s
    return _state(m)


def group_settings_greedy(settings: Iterable[InitObsSetting]) -> dict:
    """{max setting: [settings measurable with it]}, filled greedily in input order; a group that
    takes a new setting moves to the end under its new max setting
    (cirq-core/cirq/work/observable_grouping.py:27-82)."""
    grouped: dict = {}
    for setting in settings:
        for max_setting, simul in grouped.items():
            trial = [*simul, setting]
            st = _max_weight_state(s.init_state for s in trial)
            obs = _max_weight_observable(s.observable for s in trial)
            if st is not None and obs is not None:
                del grouped[max_setting]
                grouped[InitObsSetting(st, obs)] = trial
                break
        else:
            grouped[InitObsSetting(setting.init_state, setting.observable.with_coefficient(1.0))] = [setting]
    return grouped


def _pad_setting(max_setting: InitObsSetting, qubits: Sequence) -> InitObsSetting:
    """Z on every qubit the max observable skips, |0> on every qubit the state skips
    (observable_measurement.py:208-230)."""
    obs = dict(max_setting.observable.items())
    for q in qubits:
        obs.setdefault(q, Z)
    st = max_setting.init_map
    for q in qubits:
        st.setdefault(q, KET_ZERO)
    return InitObsSetting(init_state=_state(st), observable=PauliString(qubit_pauli_map=obs))


def _params_for_setting(setting: InitObsSetting, flips, qubits: Sequence, needs_init_layer: bool) -> dict:
    """Basis-change (and initialisation) exponents of every qubit (observable_measurement.py:168-205)."""
    setting = _pad_setting(setting, qubits)
    params = {}
    init = setting.init_map
    for q, flip in itertools.zip_longest(qubits, flips):
        if q is None or flip is None:
            raise ValueError("`qubits` and `flips` must be equal length")
        xf, yf = _OBS_TO_PARAM_VAL[setting.observable[q]._index, bool(flip)]
        params[f"{q}-Xf"] = xf
        params[f"{q}-Yf"] = yf
        if needs_init_layer:
            xi, yi = _STATE_TO_PARAM_VAL[init[q]]
            params[f"{q}-Xi"] = xi
            params[f"{q}-Yi"] = yi
    return params


def _with_parameterized_layers(circuit: Circuit, qubits: Sequence, needs_init_layer: bool) -> Circuit:
    """circuit flanked by X**s, Y**s layers (initialisation before, basis change after) and a
    measurement of every qubit under key 'z' (observable_measurement.py:55-80)."""
    moments = []
    if needs_init_layer:
        moments.append(Moment(*[XPowGate(exponent=sympy.Symbol(f"{q}-Xi")).on(q) for q in qubits]))
        moments.append(Moment(*[YPowGate(exponent=sympy.Symbol(f"{q}-Yi")).on(q) for q in qubits]))
    moments += list(circuit)
    moments.append(Moment(*[XPowGate(exponent=sympy.Symbol(f"{q}-Xf")).on(q) for q in qubits]))
    moments.append(Moment(*[YPowGate(exponent=sympy.Symbol(f"{q}-Yf")).on(q) for q in qubits]))
    moments.append(Moment(measure(*qubits, key="z")))
    return Circuit.from_moments(*moments)


@dataclasses.dataclass(frozen=True)
class _MeasurementSpec:
    max_setting: InitObsSetting
    circuit_params: tuple  # ((name, value), ...)

    @property
    def params(self) -> dict:
        return dict(self.circuit_params)


def _real_coefficient(obs: PauliString, atol: float) -> float:
    c = obs.coefficient
    if isinstance(c, sympy.Basic) or not np.isclose(complex(c).imag, 0, atol=atol):
        raise ValueError(f"{obs} has a complex or symbolic coefficient.")
    return complex(c).real


def _stats(bitstrings: np.ndarray, qubit_to_index: Mapping, obs: PauliString, atol: float) -> tuple[float, float]:
    """Mean of coefficient * prod (1 - 2 b_q) and its squared standard error (ddof = 1)
    (observable_measurement_data.py:47-86)."""
    idx = [qubit_to_index[q] for q in obs.keys()]
    signs = 1 - 2 * np.asarray(bitstrings[:, idx], dtype=np.int8)
    vals = _real_coefficient(obs, atol) * np.prod(signs, axis=1)
    return float(np.mean(vals)), float(np.var(vals, ddof=1) / len(vals))


@dataclasses.dataclass(frozen=True)
class ObservableMeasuredResult:
    setting: InitObsSetting
    mean: float
    variance: float
    repetitions: int
    circuit_params: Mapping

    @property
    def observable(self) -> PauliString:
        return self.setting.observable

    @property
    def stddev(self) -> float:
        return float(np.sqrt(self.variance))


class BitstringAccumulator:
    """The bitstrings measured for one measurement spec and the statistics of each setting it
    serves (observable_measurement_data.py:170-530, without readout calibration)."""

    def __init__(self, meas_spec: _MeasurementSpec, simul_settings: list, qubit_to_index: Mapping):
        self.meas_spec = meas_spec
        self.simul_settings = list(simul_settings)
        self.qubit_to_index = dict(qubit_to_index)
        self.bitstrings = np.zeros((0, len(qubit_to_index)), dtype=np.uint8)
        self.chunksizes = np.zeros(0, dtype=np.int64)

    def consume_results(self, bitstrings: np.ndarray):
        if bitstrings.dtype != np.uint8:
            raise ValueError("`bitstrings` should be of type np.uint8")
        self.bitstrings = np.append(self.bitstrings, bitstrings, axis=0)
        self.chunksizes = np.append(self.chunksizes, [len(bitstrings)])

    @property
    def n_repetitions(self) -> int:
        return len(self.bitstrings)

    def _check(self, setting: InitObsSetting, what: str):
        if len(self.bitstrings) == 0:
            raise ValueError("No measurements")
        if setting not in self.simul_settings:
            raise ValueError(f"{setting} is not one of the settings of this accumulator ({what})")

    def mean(self, setting: InitObsSetting, *, atol: float = 1e-8) -> float:
        self._check(setting, "mean")
        return _stats(self.bitstrings, self.qubit_to_index, setting.observable, atol)[0]

    def variance(self, setting: InitObsSetting, *, atol: float = 1e-8) -> float:
        self._check(setting, "variance")
        return _stats(self.bitstrings, self.qubit_to_index, setting.observable, atol)[1]

    def stderr(self, setting: InitObsSetting, *, atol: float = 1e-8) -> float:
        return float(np.sqrt(self.variance(setting, atol=atol)))

    def means(self, *, atol: float = 1e-8) -> np.ndarray:
        return np.asarray([self.mean(s, atol=atol) for s in self.simul_settings])

    @property
    def results(self):
        for s in self.simul_settings:
            m, v = _stats(self.bitstrings, self.qubit_to_index, s.observable, 1e-8)
            yield ObservableMeasuredResult(setting=s, mean=m, variance=v, repetitions=len(self.bitstrings),
                                           circuit_params=self.meas_spec.params)


class StoppingCriteria:
    def more_repetitions(self, accumulator: BitstringAccumulator) -> int:
        raise NotImplementedError


@dataclasses.dataclass(frozen=True)
class RepetitionsStoppingCriteria(StoppingCriteria):
    """Stop at a total number of repetitions, in chunks (observable_measurement.py:123-140)."""

    total_repetitions: int
    repetitions_per_chunk: int = 10_000

    def more_repetitions(self, accumulator: BitstringAccumulator) -> int:
        todo = self.total_repetitions - accumulator.n_repetitions
        return 0 if todo <= 0 else min(self.repetitions_per_chunk, todo)


@dataclasses.dataclass(frozen=True)
class VarianceStoppingCriteria(StoppingCriteria):
    """Stop when every setting's squared standard error is below a bound
    (observable_measurement.py:97-120)."""

    variance_bound: float
    repetitions_per_chunk: int = 10_000

    def more_repetitions(self, accumulator: BitstringAccumulator) -> int:
        if len(accumulator.bitstrings) == 0:
            return self.repetitions_per_chunk
        cov = np.array([accumulator.variance(s) for s in accumulator.simul_settings])
        return self.repetitions_per_chunk if np.any(cov > self.variance_bound) else 0


def _still_todo(specs: list, accumulators: Mapping, criteria: StoppingCriteria) -> tuple[list, int]:
    """Specs that need more repetitions and the chunk size for all of them
    (observable_measurement.py:248-297)."""
    todo, reps_set = [], set()
    for spec in specs:
        more = criteria.more_repetitions(accumulators[spec])
        if more < 0:
            raise ValueError("Stopping criteria's `more_repetitions` should return 0 or a positive number.")
        if more == 0:
            continue
        reps_set.add(more)
        todo.append(spec)
    if not todo:
        return todo, 0
    reps = next(iter(reps_set)) if len(reps_set) == 1 else max(reps_set)
    if len(reps_set) > 1:
        warnings.warn(
            "The stopping criteria specified a various numbers of repetitions to perform next. To be able to "
            f"submit as a single sweep, the largest value will be used: {reps}."
        )
    if len(todo) * reps > MAX_REPETITIONS_PER_JOB:
        old = reps
        reps = MAX_REPETITIONS_PER_JOB // len(todo)
        if reps < 10:
            raise ValueError(
                "You have requested too many parameter settings to batch your job effectively. "
                "Consider fewer sweeps or manually splitting sweeps into multiple jobs."
            )
        warnings.warn(
            f"The number of requested sweep parameters is high. To avoid a batched job with more than "
            f"{MAX_REPETITIONS_PER_JOB} shots, the number of shots per call to run_sweep (per parameter value) "
            f"will be throttled from {old} to {reps}."
        )
    return todo, reps


def measure_grouped_settings(circuit: Circuit, grouped_settings: Mapping, sampler, stopping_criteria: StoppingCriteria,
                             *, readout_symmetrization: bool = False, circuit_sweep=None) -> list[BitstringAccumulator]:
    """Measures every group for every circuit parameter point until the stopping criteria are met
    (observable_measurement.py:473-588; readout calibration is not supported)."""
    qubits = sorted({q for ms in grouped_settings for q, _ in ms.init_state})
    qubit_to_index = {q: i for i, q in enumerate(qubits)}
    needs_init = any(s != KET_ZERO for ms in grouped_settings for _, s in ms.init_state)
    program = _with_parameterized_layers(circuit, qubits, needs_init)
    accumulators: dict = {}
    todo: list = []
    for max_setting, resolver in itertools.product(list(grouped_settings), list(to_resolvers(circuit_sweep))):
        spec = _MeasurementSpec(max_setting, tuple(resolver.param_dict.items()))
        accumulators[spec] = BitstringAccumulator(spec, grouped_settings[max_setting], qubit_to_index)
        todo.append(spec)
    n = len(qubits)
    while True:
        todo, reps = _still_todo(todo, accumulators, stopping_criteria)
        if not todo:
            break
        flippy = []
        for spec in todo:
            flippy.append((spec, np.zeros(n, dtype=bool)))
            if readout_symmetrization:
                flippy.append((spec, np.ones(n, dtype=bool)))
        if readout_symmetrization:
            reps //= 2
        resolvers = []
        for spec, flips in flippy:
            params = _params_for_setting(spec.max_setting, flips, qubits, needs_init)
            params.update(spec.params)
            resolvers.append(ParamResolver(params))
        results = sampler.run_sweep(program, params=resolvers, repetitions=reps)
        assert len(results) == len(flippy), "Not as many results received as sweeps requested!"
        for (spec, flips), res in zip(flippy, results):
            bits = np.logical_xor(flips, res.measurements["z"])
            accumulators[spec].consume_results(bits.astype(np.uint8, casting="safe"))
    return list(accumulators.values())


def _all_qubits(circuit: Circuit, observables: Iterable[PauliString]) -> list:
    qs = set()
    for obs in observables:
        qs |= set(obs.qubits)
    qs |= set(circuit.all_qubits())
    return sorted(qs)


def flatten_grouped_results(grouped: list[BitstringAccumulator]) -> list[ObservableMeasuredResult]:
    return [r for acc in grouped for r in acc.results]


def measure_observables(circuit: Circuit, observables: Iterable[PauliString], sampler,
                        stopping_criteria: StoppingCriteria, *, readout_symmetrization: bool = False,
                        circuit_sweep=None,
                        grouper: str | Callable = "greedy") -> list[ObservableMeasuredResult]:
    """One ObservableMeasuredResult per input PauliString (observable_measurement.py:614-673)."""
    observables = list(observables)
    qubits = _all_qubits(circuit, observables)
    settings = list(observables_to_settings(observables, qubits))
    if isinstance(grouper, str):
        if grouper.lower() != "greedy":
            raise ValueError(f"Unknown grouping function {grouper}")
        grouper = group_settings_greedy
    acc = measure_grouped_settings(circuit, grouper(settings), sampler, stopping_criteria,
                                   readout_symmetrization=readout_symmetrization, circuit_sweep=circuit_sweep)
    return flatten_grouped_results(acc)


def _param_key(items) -> frozenset:
    """Fixed-precision hashable form of a parameter assignment (observable_settings.py:129-160)."""
    out = []
    for k, v in items:
        if not isinstance(k, str):
            continue
        if isinstance(v, sympy.Basic):
            raise ValueError(f"Cannot convert {v} to fixed precision in observable settings")
        v = complex(v)
        out.append((k, (int(v.real * 1e7), int(v.imag * 1e7)) if v.imag else int(v.real * 1e7)))
    return frozenset(out)


def sample_expectation_values(sampler, program: Circuit, observables, *, num_samples: int, params=None,
                              permit_terminal_measurements: bool = False) -> list[list[float]]:
    """results[sweep point][observable]: sampled estimates of <observable> (cirq-core/cirq/work/
    sampler.py:311-400).  As there, each PauliString is credited to the last observable (sum)
    that contains it: a string shared by two sums adds both of its estimates to the later sum."""
    if num_samples <= 0:
        raise ValueError(f"Expectation values require at least one sample. Received: {num_samples}.")
    if not observables:
        raise ValueError("At least one observable must be provided.")
    if not permit_terminal_measurements and program.are_any_measurements_terminal():
        raise ValueError(
            "Provided circuit has terminal measurements, which may skew expectation values. If this is "
            "intentional, set permit_terminal_measurements=True."
        )
    sums = [PauliSum.wrap(o) for o in observables] if isinstance(observables, list) else [PauliSum.wrap(observables)]
    flat: list[PauliString] = []
    to_sum: dict = {}
    for i, s in enumerate(sums):
        for ps in s:
            flat.append(ps)
            to_sum[ps] = i
    points = [r.param_dict for r in to_resolvers(params)]
    point_of = {_param_key(p.items()): i for i, p in enumerate(points)}
    results = measure_observables(program, flat, sampler, RepetitionsStoppingCriteria(total_repetitions=num_samples),
                                  circuit_sweep=params)
    nested = [[0.0] * len(sums) for _ in points]
    for res in results:
        nested[point_of[_param_key(res.circuit_params.items())]][to_sum[res.setting.observable]] += res.mean
    return nested


__all__ = [
    "BitstringAccumulator", "InitObsSetting", "ObservableMeasuredResult", "RepetitionsStoppingCriteria",
    "VarianceStoppingCriteria", "flatten_grouped_results", "group_settings_greedy", "measure_grouped_settings",
    "measure_observables", "observables_to_settings", "sample_expectation_values", "zeros_state",
]
