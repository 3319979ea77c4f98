"""Parameter resolution, sweeps and results (cirq-core/cirq/study/).

ParamResolver.value_of follows cirq-core/cirq/study/resolver.py:108-300 for numbers, symbols,
names and sympy expressions (evaluated by substitution); sweeps follow
cirq-core/cirq/study/sweeps.py and to_resolvers follows cirq-core/cirq/study/sweepable.py:43-112.
"""

from __future__ import annotations

import collections
import itertools
import numbers
from typing import Any, Iterable, Iterator, Mapping, Sequence

import numpy as np
import sympy


class ParamResolver:
    """Maps symbols (or their names) to values."""

    def __init__(self, param_dict: Mapping | "ParamResolver" | None = None):
        if isinstance(param_dict, ParamResolver):
            param_dict = param_dict.param_dict
        self._param_dict = dict(param_dict or {})
        self._by_name = {}
        for k, v in self._param_dict.items():
            name = k.name if isinstance(k, sympy.Symbol) else str(k)
            self._by_name[name] = v

    @property
    def param_dict(self) -> dict:
        return self._param_dict

    def value_of(self, value, recursive: bool = True):
        if isinstance(value, (numbers.Number, np.number)) and not isinstance(value, sympy.Basic):
            return value
        if isinstance(value, str):
            if value in self._by_name:
                return self._resolve_raw(self._by_name[value], recursive)
            return sympy.Symbol(value)
        if isinstance(value, sympy.Symbol):
            if value.name in self._by_name:
                return self._resolve_raw(self._by_name[value.name], recursive)
            if value in self._param_dict:
                return self._resolve_raw(self._param_dict[value], recursive)
            return value
        if isinstance(value, sympy.Basic):
            if not value.free_symbols:
                v = complex(value)
                return v.real if v.imag == 0 else v
            subs = {}
            for s in value.free_symbols:
                r = self.value_of(s, recursive)
                subs[s] = r
            out = value.subs(subs)
            if isinstance(out, sympy.Basic) and not out.free_symbols:
                v = complex(out)
                return v.real if v.imag == 0 else v
            return out
        return value

    def _resolve_raw(self, v, recursive):
        if recursive and isinstance(v, (sympy.Basic, str)) and not isinstance(v, numbers.Number):
            return self.value_of(v, recursive)
        return v

    def __getitem__(self, key):
        return self.value_of(key)

    def __contains__(self, key):
        name = key.name if isinstance(key, sympy.Symbol) else str(key)
        return name in self._by_name

    def __bool__(self):
        return bool(self._param_dict)

    def __iter__(self):
        return iter(self._param_dict)

    def __len__(self):
        return len(self._param_dict)

    def __eq__(self, other):
        if not isinstance(other, ParamResolver):
            return NotImplemented
        return self._by_name == other._by_name

    def __hash__(self):
        return hash(tuple(sorted((k, str(v)) for k, v in self._by_name.items())))

    def __repr__(self):
        return f"cirq.ParamResolver({self._param_dict!r})"


# ---- sweeps --------------------------------------------------------------------------------------
class Sweep:
    """A sequence of parameter assignments."""

    @property
    def keys(self) -> list:
        raise NotImplementedError

    def __len__(self) -> int:
        raise NotImplementedError

    def param_tuples(self) -> Iterator[tuple]:
        raise NotImplementedError

    def __iter__(self) -> Iterator[ParamResolver]:
        for pt in self.param_tuples():
            yield ParamResolver(dict(pt))

    def __mul__(self, other: "Sweep") -> "Product":
        return Product(self, other)

    def __add__(self, other: "Sweep") -> "Zip":
        return Zip(self, other)

    def __getitem__(self, i):
        return list(self)[i]


class _UnitSweep(Sweep):
    keys = []

    def __len__(self):
        return 1

    def param_tuples(self):
        yield ()


UnitSweep = _UnitSweep()


def _key_name(k) -> str:
    return k.name if isinstance(k, sympy.Symbol) else str(k)


class Points(Sweep):
    def __init__(self, key, points: Sequence):
        self.key = key
        self.points = list(points)

    @property
    def keys(self):
        return [self.key]

    def __len__(self):
        return len(self.points)

    def param_tuples(self):
        for p in self.points:
            yield ((_key_name(self.key), p),)


class Linspace(Sweep):
    def __init__(self, key, start: float, stop: float, length: int):
        self.key = key
        self.start, self.stop, self.length = start, stop, int(length)

    @property
    def keys(self):
        return [self.key]

    def __len__(self):
        return self.length

    def param_tuples(self):
        if self.length == 1:
            yield ((_key_name(self.key), self.start),)
            return
        for i in range(self.length):
            p = i / (self.length - 1)
            yield ((_key_name(self.key), self.start * (1 - p) + self.stop * p),)


class ListSweep(Sweep):
    def __init__(self, resolver_list: Iterable):
        self.resolver_list = [r if isinstance(r, ParamResolver) else ParamResolver(r) for r in resolver_list]

    @property
    def keys(self):
        ks = []
        for r in self.resolver_list:
            for k in r.param_dict:
                if _key_name(k) not in ks:
                    ks.append(_key_name(k))
        return ks

    def __len__(self):
        return len(self.resolver_list)

    def param_tuples(self):
        for r in self.resolver_list:
            yield tuple((_key_name(k), v) for k, v in r.param_dict.items())


class Zip(Sweep):
    def __init__(self, *sweeps: Sweep):
        self.sweeps = sweeps

    @property
    def keys(self):
        return [k for s in self.sweeps for k in s.keys]

    def __len__(self):
        return min((len(s) for s in self.sweeps), default=0)

    def param_tuples(self):
        iters = [s.param_tuples() for s in self.sweeps]
        for vals in zip(*iters):
            yield sum(vals, ())


class Product(Sweep):
    def __init__(self, *factors: Sweep):
        self.factors = factors

    @property
    def keys(self):
        return [k for s in self.factors for k in s.keys]

    def __len__(self):
        n = 1
        for f in self.factors:
            n *= len(f)
        return n

    def param_tuples(self):
        for combo in itertools.product(*[list(f.param_tuples()) for f in self.factors]):
            yield sum(combo, ())


def to_resolvers(sweepable) -> Iterator[ParamResolver]:
    """cirq.to_resolvers (cirq-core/cirq/study/sweepable.py:43-46)."""
    if sweepable is None:
        yield ParamResolver({})
    elif isinstance(sweepable, ParamResolver):
        yield sweepable
    elif isinstance(sweepable, Sweep):
        yield from sweepable
    elif isinstance(sweepable, dict):
        yield ParamResolver(sweepable)
    elif isinstance(sweepable, Iterable) and not isinstance(sweepable, str):
        for item in sweepable:
            yield from to_resolvers(item)
    else:
        raise TypeError(f"Unrecognized sweepable type: {type(sweepable)}.")


# ---- results ------------------------------------------------------------------------------------
class ResultDict:
    """Measurement records of one run (cirq-core/cirq/study/result.py:378).

    ``records[key]`` has shape (repetitions, instances, n_qubits); ``measurements[key]`` is the
    last instance, shape (repetitions, n_qubits).
    """

    def __init__(self, *, params: ParamResolver | None = None, measurements=None, records=None):
        self._params = params or ParamResolver({})
        if records is None:
            records = {k: np.asarray(v)[:, None, :] for k, v in (measurements or {}).items()}
        self._records = dict(records)

    @property
    def params(self) -> ParamResolver:
        return self._params

    @property
    def records(self) -> dict:
        return self._records

    @property
    def measurements(self) -> dict:
        return {k: v[:, -1, :] for k, v in self._records.items()}

    @property
    def repetitions(self) -> int:
        if not self._records:
            return 0
        return next(iter(self._records.values())).shape[0]

    def histogram(self, *, key, fold_func=None) -> collections.Counter:
        bits = self.measurements[key]
        if fold_func is None:
            w = 1 << np.arange(bits.shape[1] - 1, -1, -1, dtype=np.int64)
            vals = (bits.astype(np.int64) * w).sum(axis=1)
            return collections.Counter(vals.tolist())
        return collections.Counter(fold_func(b) for b in bits)

    def multi_measurement_histogram(self, *, keys, fold_func=None) -> collections.Counter:
        keys = list(keys)
        ms = [self.measurements[k] for k in keys]
        out = collections.Counter()
        for r in range(self.repetitions):
            samples = tuple(m[r] for m in ms)
            if fold_func is None:
                key = tuple(int("".join(str(int(b)) for b in s) or "0", 2) for s in samples)
            else:
                key = fold_func(samples)
            out[key] += 1
        return out

    def __repr__(self):
        return f"cirq.ResultDict(params={self._params!r}, records={self._records!r})"
