"""Per-circuit choice of tile shape and fusion settings, by timing candidate programs on the GPU.

The pass compiler has a handful of knobs whose best values depend on the circuit and the device:
the tile width L (more gates per pass vs. fewer workgroups in flight), the coalescing bits (how
many low flat bits every tile keeps), the widest fused block and the number of outer bits a fused
op may depend on, how the planner weighs MFMA blocks against VALU gates, and whether small blocks
left at the end of a pass share MFMA blocks.  `Autotuner`
compiles candidates for a lowered circuit, runs each on a scratch state with HIP events around a
few repetitions, walks the knobs one at a time from the default (coordinate descent), and caches
the fastest per circuit structure (optionally in a JSON file, so a later process reuses it).

Cirq has no counterpart: its NumPy path has nothing to tune.
"""

from __future__ import annotations

import dataclasses
import hashlib
import json
import os

import torch

from . import engine
from .tape import COSTS, DEFAULT_KMAX, MAX_VARIANT_BITS, SPECIALIZED_COSTS, LOp, compile_ops


@dataclasses.dataclass(frozen=True)
class TuneConfig:
    tile_bits: int = 12
    coalesce_bits: int = 4
    kmax: int = 5  # widest fused block
    variant_bits: int = MAX_VARIANT_BITS  # outer bits a fused op may depend on
    mfma_weight: float = 1.0  # multiplies the MFMA entries of the planner's cost table
    blocks: bool = True
    kron_merge: bool = True  # small blocks left at the end of a pass share MFMA blocks

    def compile_kwargs(self, n_bits: int, specialized: bool = False) -> dict:
        """compile_ops keyword arguments; specialized: the program will run as specialised kernels
        (the planner's cost table takes SPECIALIZED_COSTS)."""
        L = min(self.tile_bits, n_bits)
        costs = dict(COSTS, **SPECIALIZED_COSTS) if specialized else dict(COSTS)
        costs["mfma4"] *= self.mfma_weight
        costs["mfma5"] *= self.mfma_weight
        return dict(L=L, coalesce=min(self.coalesce_bits, L), blocks=self.blocks, kmax=self.kmax, costs=costs,
                    variant_bits=self.variant_bits, kron_merge=self.kron_merge)


def default_config(n_bits: int, dm: bool = False) -> TuneConfig:
    """The settings the simulators use without tuning."""
    L = 12 if n_bits >= 20 else (10 if n_bits >= 14 else n_bits)
    return TuneConfig(tile_bits=min(L, n_bits), coalesce_bits=min(4, L), kmax=DEFAULT_KMAX[bool(dm)])


def _axes(n_bits: int, base: TuneConfig) -> list[tuple[str, list]]:
    """The knobs and the values tried for each (the base value first)."""
    Ls = [L for L in (base.tile_bits, base.tile_bits - 1, base.tile_bits + 1) if 8 <= L <= min(13, n_bits)]
    return [
        ("tile_bits", Ls or [base.tile_bits]),
        ("mfma_weight", [base.mfma_weight, 0.7, 0.85, 1.2]),
        ("variant_bits", [base.variant_bits, 2]),
        ("kmax", [base.kmax, 9 - base.kmax]),
        ("coalesce_bits", [base.coalesce_bits, 3, 5]),
        ("kron_merge", [base.kron_merge, not base.kron_merge]),
        ("blocks", [base.blocks, not base.blocks]),
    ]


def structure_key(ops: list[LOp], n_bits: int) -> str:
    """A fingerprint of what the compiler sees: op kinds, bits and which ops are parameterised
    (not the parameter values, which do not change the program's shape)."""
    h = hashlib.sha1(str(n_bits).encode())
    for op in ops:
        h.update(repr((op.kind, op.targets, op.controls, op.pref is None)).encode())
    return h.hexdigest()


class Autotuner:
    """Chooses a TuneConfig per circuit structure.

    path: optional JSON file of tuned configurations (read at start, rewritten after each tuning).
    reps: timed repetitions per candidate (after one warm-up run).
    """

    def __init__(self, path: str | None = None, reps: int = 2, min_gain: float = 0.02):
        self.path = path
        self.reps = max(1, int(reps))
        self.min_gain = float(min_gain)
        self.cache: dict[str, TuneConfig] = {}
        self.log: dict[str, list] = {}  # key -> [(config, ms), ...] of the last tuning
        if path and os.path.exists(path):
            with open(path) as f:
                for k, v in json.load(f).items():
                    self.cache[k] = TuneConfig(**v)

    def _save(self):
        if not self.path:
            return
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump({k: dataclasses.asdict(v) for k, v in self.cache.items()}, f, indent=1)
        os.replace(tmp, self.path)

    def time_program(self, ops: list[LOp], n_bits: int, cfg: TuneConfig, state: torch.Tensor,
                     basis: int | None = 0) -> float:
        """Average ms of one run of the program compiled with cfg (HIP events on torch's stream),
        from |basis> as simulations from a computational basis state run (the passes then skip the
        tiles the state cannot reach yet); basis None: from the state in the buffer."""
        prog = engine.DeviceProgram(compile_ops(ops, n_bits, **cfg.compile_kwargs(n_bits)), state.device)
        prog.run(state, basis=basis)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(self.reps):
            prog.run(state, basis=basis)
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1) / self.reps

    def tune(self, ops: list[LOp], n_bits: int, base: TuneConfig | None = None, dm: bool = False) -> TuneConfig:
        """Coordinate descent from `base` (default_config): each knob in turn takes the value
        with the lowest time, keeping a change only if it saves more than min_gain."""
        device = engine.require_gpu()
        base = base or default_config(n_bits, dm)
        state = engine.new_state(n_bits, device)
        timings: dict[TuneConfig, float] = {}

        def t(cfg: TuneConfig) -> float:
            if cfg not in timings:
                timings[cfg] = self.time_program(ops, n_bits, cfg, state)
            return timings[cfg]

        best = base
        best_t = t(base)
        for name, values in _axes(n_bits, base):
            for v in values[1:]:
                cand = dataclasses.replace(best, **{name: v})
                tc = t(cand)
                if tc < best_t * (1 - self.min_gain):
                    best, best_t = cand, tc
        del state
        key = structure_key(ops, n_bits)
        self.cache[key] = best
        self.log[key] = sorted(((c, ms) for c, ms in timings.items()), key=lambda x: x[1])
        self._save()
        return best

    def config_for(self, ops: list[LOp], n_bits: int, dm: bool = False) -> TuneConfig:
        key = structure_key(ops, n_bits)
        hit = self.cache.get(key)
        return hit if hit is not None else self.tune(ops, n_bits, dm=dm)


_default: Autotuner | None = None


def default_autotuner() -> Autotuner:
    """Process-wide tuner; CIRQ_AMD_TUNE_CACHE names a JSON file to keep results across runs."""
    global _default
    if _default is None:
        _default = Autotuner(os.environ.get("CIRQ_AMD_TUNE_CACHE"))
    return _default


def tuned_compile_kwargs(ops: list[LOp], n_bits: int, tuner: Autotuner | None = None, dm: bool = False) -> dict:
    return (tuner or default_autotuner()).config_for(ops, n_bits, dm).compile_kwargs(n_bits)


__all__ = ["Autotuner", "TuneConfig", "default_config", "default_autotuner", "structure_key", "tuned_compile_kwargs"]
