"""Batched trajectories: many repetitions of a circuit with mid-circuit measurements, or with
channels on a state vector, simulated at once.

Cirq repeats such a circuit's general suffix once per repetition on a copy of the prefix state
(cirq-core/cirq/sim/simulator_base.py:248-273).  Here R repetitions are R copies of the prefix
state stored one after the other, i.e. one state of n_bits + log2(R) bits whose top bits number
the copy.  A deterministic stretch of the suffix is one pass program over all of them (no op
touches the copy bits); a measurement, mixture or Kraus channel is one batched kernel that gives
every copy its own outcome (`cqs_marginal_probs_batch`, `cqs_draw_rows`, `cqs_collapse_batch`,
`cqs_apply_select_batch`, `cqs_kraus_weights_batch`).

Every stochastic op consumes exactly one uniform of the simulator's RandomState per repetition,
as in Cirq: `prng.choice(len(p), p=p)` for a measurement
(cirq-core/cirq/sim/state_vector.py:235-322) and for a mixture
(cirq-core/cirq/sim/state_vector_simulation_state.py:183-203), `prng.random()` for a Kraus
channel (:205-257).  The uniforms are drawn up front in Cirq's order (repetition by repetition,
op by op within one) and each copy uses its own row, with NumPy's choice arithmetic restated: on
the GPU for measurements (`cqs_draw_rows`), below for mixtures (whose probabilities are the same
for every copy); so the records equal those of the one-copy-at-a-time loop except where a
probability computed by another kernel rounds differently at a draw's boundary.
"""

from __future__ import annotations


import numpy as np
import torch

from . import engine
from .tape import DEFAULT_KMAX, Measurement, ParamEvaluator, _is_measurement, compile_ops, fill_param_mats, kraus_of

MAX_BATCH_CHANNEL_QUBITS = 3
MAX_KRAUS = 16


# ---------------------------------------------------------------- NumPy's draws, vectorised
def choice_rows(p: np.ndarray, u: np.ndarray) -> np.ndarray:
    """Row r: RandomState.choice(len(p[r]), p=p[r]) given that its uniform is u[r]: the cumulative
    sum in float64 normalised by its last entry, then searchsorted(..., 'right')."""
    cdf = np.cumsum(np.asarray(p, dtype=np.float64), axis=-1)
    cdf = cdf / cdf[..., -1:]
    return np.sum(cdf <= np.asarray(u, dtype=np.float64)[:, None], axis=-1).astype(np.int64)


def kraus_draws(weights: np.ndarray, u: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Kraus operator index and its weight per row, as the walk of
    cirq-core/cirq/sim/state_vector_simulation_state.py:224-250 picks them: p = u, then for each
    operator in order p -= ||K_i psi||^2 until p < 0; the heaviest operator seen so far (first of
    equals) if the walk ends with p >= 0 or on a zero weight."""
    R, M = weights.shape
    p = np.asarray(u, dtype=np.float64).copy()
    done = np.zeros(R, dtype=bool)
    sel = np.zeros(R, dtype=np.int64)
    last_w = np.zeros(R)
    fb_idx = np.zeros(R, dtype=np.int64)
    fb_w = np.zeros(R)
    for i in range(M):
        act = ~done
        w = weights[:, i]
        better = act & (w > fb_w)
        fb_idx[better] = i
        fb_w[better] = w[better]
        p[act] -= w[act]
        last_w[act] = w[act]
        hit = act & (p < 0)
        sel[hit] = i
        done |= hit
    fallback = ~done | (last_w == 0)
    return np.where(fallback, fb_idx, sel), np.where(fallback, fb_w, last_w)


# ---------------------------------------------------------------- the batched runner
def _next_pow2(x: int) -> int:
    return 1 << max(0, (int(x) - 1).bit_length())


class TrajectoryBatch:
    """Runs a list of actions (('program', compiled), ('op', stochastic op), as the simulator
    builds them) on many copies of a state.

    qubits: the simulator's qubit order; dm: density-matrix copies (then only measurements are
    stochastic); tile_params: the simulator's (n_bits) -> (L, coalesce); max_copies: copies per
    launch (a power of two), or None to size batches by free device memory."""

    def __init__(self, qubits, actions, dm: bool, device, tile_params, fuse: bool, blocks: bool,
                 ws: engine.Workspace, max_copies: int | None = None):
        self.qubits = tuple(qubits)
        self.n = len(self.qubits)
        self.dm = dm
        self.n_bits = 2 * self.n if dm else self.n
        self.actions = actions
        self.device = device
        self.tile_params = tile_params
        self.fuse = fuse
        self.blocks = blocks
        self.ws = ws
        self.max_copies = max_copies
        self.index = {q: i for i, q in enumerate(self.qubits)}
        self._programs: dict = {}
        self.n_draws = sum(1 for kind, _ in actions if kind == "op")
        self._channels: dict = {}  # stochastic op index -> its matrices on the device
        self._mix_mats: dict = {}  # stochastic op index -> its matrices (host complex64, flat)
        self._plan = None  # the actions as steps, with runs of mixtures merged (built on first use)

    @staticmethod
    def supported(actions, dm: bool) -> bool:
        """Every stochastic op fits the batched kernels (channels on <= 3 qubits, <= 16 Kraus
        operators); measurements always do."""
        for kind, op in actions:
            if kind != "op" or dm or _is_measurement(op):
                continue
            if len(op.qubits) > MAX_BATCH_CHANNEL_QUBITS:
                return False
            obj = getattr(op, "gate", None) or op
            mix = getattr(obj, "_mixture_", None)
            mixture = mix() if mix is not None else None
            if mixture is None or mixture is NotImplemented:
                if len(kraus_of(obj)) > MAX_KRAUS:
                    return False
        return True

    def copies_per_batch(self, repetitions: int) -> int:
        per = 8 << self.n_bits
        if self.max_copies is not None:
            cap = max(1, int(self.max_copies))
        else:
            free, _ = torch.cuda.mem_get_info(self.device)
            cap = max(1, int(free // 4 // per))
        cap = min(cap, 32768, _next_pow2(repetitions))
        return 1 << (cap.bit_length() - 1)

    def _program(self, comp, copy_bits: int):
        key = (id(comp), copy_bits)
        hit = self._programs.get(key)
        if hit is None:
            nb = comp.n_bits + copy_bits
            L, c = self.tile_params(nb)
            packed = compile_ops(comp.lowering.ops, nb, L=L, coalesce=c, fuse=self.fuse, blocks=self.blocks,
                                 kmax=DEFAULT_KMAX[bool(self.dm)])
            hit = (packed, engine.DeviceProgram(packed, self.device), ParamEvaluator(comp.lowering.param_refs))
            self._programs[key] = hit
        return hit

    def run(self, state0: torch.Tensor, resolver, repetitions: int, prng: np.random.RandomState) -> list:
        """Per repetition, the list of (key, bits) measurement records in circuit order."""
        D = self.n_draws
        U = prng.random_sample(repetitions * D).reshape(repetitions, D) if D else np.zeros((repetitions, 0))
        out: list = [[] for _ in range(repetitions)]
        step = self.copies_per_batch(repetitions)
        for start in range(0, repetitions, step):
            valid = min(step, repetitions - start)
            R = _next_pow2(valid)
            u = U[start:start + valid]
            if R > valid:  # padding copies repeat the first valid copy's draws (and so its state)
                u = np.concatenate([u, np.repeat(u[:1], R - valid, axis=0)])
            recs = self._run_batch(state0, resolver, R, u)
            for key, bits in recs:
                for r in range(valid):
                    out[start + r].append((key, bits[r]))
        return out

    def _channel_of(self, j: int, op):
        """(targets, device matrices, count, mixture probabilities or None) of stochastic op j."""
        hit = self._channels.get(j)
        if hit is None:
            targets = [self.n - 1 - self.index[q] for q in reversed(op.qubits)]
            obj = getattr(op, "gate", None) or op
            mix = getattr(obj, "_mixture_", None)
            mixture = mix() if mix is not None else None
            if mixture is not None and mixture is not NotImplemented:
                probabilities, mats = zip(*mixture)
                probs = np.asarray(probabilities, dtype=np.float64)
            else:
                mats, probs = kraus_of(obj), None
            mats = np.stack([np.asarray(x, dtype=np.complex128) for x in mats])
            m_dev, n_mats = engine.device_mats(mats, targets, self.device)
            hit = (targets, m_dev, n_mats, probs)
            self._channels[j] = hit
            self._mix_mats[j] = mats.astype(np.complex64).reshape(-1)
        return hit

    def _steps(self):
        """The actions as steps: ('program', compiled), ('measure', j, op), ('kraus', [j, ...]) for
        Kraus channels in a row (one-qubit ones chained), ('mix', j, op) for a mixture on 3 qubits,
        and ('run', [j, ...]) for consecutive mixtures on 1-2 qubits whose targets span at most 4
        bits (one sweep of the copies for all of them).
        Mixture j's matrices sit in one device buffer at an offset; row j of the selection table
        holds every copy's choice for it."""
        if self._plan is not None:
            return self._plan
        steps, run, run_bits = [], [], set()
        offsets, bufs, total = {}, [], 0

        def flush():
            nonlocal run, run_bits
            if run:
                steps.append(("run", list(run)))
            run, run_bits = [], set()

        j = 0
        for kind, a in self.actions:
            if kind == "program":
                flush()
                steps.append(("program", a))
            elif kind == "op":
                if _is_measurement(a) or self.dm:
                    flush()
                    steps.append(("measure", j, a))
                else:
                    targets, _, n_mats, probs = self._channel_of(j, a)
                    if probs is None:
                        flush()
                        prev = steps[-1] if steps else None
                        if (len(targets) == 1 and n_mats <= MAX_KRAUS and prev is not None and prev[0] == "kraus"
                                and len(self._channels[prev[1][-1]][0]) == 1):
                            prev[1].append(j)  # one-qubit channels in a row: one sweep per channel
                        else:
                            steps.append(("kraus", [j]))
                    elif len(targets) > 2:
                        flush()
                        steps.append(("mix", j, a))
                    else:
                        if len(run) >= 16 or len(run_bits | set(targets)) > 4:
                            flush()
                        run.append(j)
                        run_bits |= set(targets)
                        d = 1 << len(targets)
                        offsets[j] = total
                        total += n_mats * d * d
                        bufs.append(self._mix_mats[j])
                j += 1
        flush()
        self._run_mats = engine._dev_array(np.concatenate(bufs).view(np.float32), self.device) if bufs else None
        self._run_offsets = offsets
        self._plan = steps
        return steps

    def _run_batch(self, state0: torch.Tensor, resolver, R: int, u: np.ndarray) -> list:
        copy_bits = R.bit_length() - 1
        states = state0.repeat(R)
        steps = self._steps()
        # a mixture's draw does not depend on the state: every copy's choice for every mixture is
        # known now and goes to the device in one copy (row j: stochastic op j)
        D = u.shape[1]
        table = np.full((max(D, 1), R), -1, dtype=np.int32)
        for j, (targets, _, n_mats, probs) in self._channels.items():
            if probs is not None:
                table[j] = choice_rows(np.broadcast_to(probs, (R, len(probs))), u[:, j])
        table_dev = engine._dev_array(table, self.device)
        ones = torch.ones(R, dtype=torch.float32, device=self.device)
        u_dev = engine._dev_array(np.ascontiguousarray(u.T), self.device) if D else None  # row j: op j
        recs: list = []
        for step in steps:
            kind = step[0]
            if kind == "program":
                packed, prog, ev = self._program(step[1], copy_bits)
                if packed.param_entries:
                    prog.set_mats(fill_param_mats(packed, ev, ev.symbol_values([resolver]))[0])
                prog.run(states)
            elif kind == "measure":
                recs.append(self._measure(states, R, step[2], u[:, step[1]]))
            elif kind == "kraus":
                self._kraus_chain(states, R, step[1], u_dev)
            elif kind == "mix":
                targets, m_dev, n_mats, _ = self._channel_of(step[1], step[2])
                engine.apply_select_batch_dev(states, self.n_bits, R, targets, m_dev, n_mats, table_dev[step[1]], ones)
            else:
                ops = []
                for j in step[1]:
                    targets, _, n_mats, _ = self._channels[j]
                    ops.append((len(targets), targets, self._run_offsets[j], n_mats, j))
                engine.select_run_batch(states, self.n_bits, R, ops, self._run_mats, table_dev)
        return recs

    def _measure(self, states, R: int, op, u: np.ndarray):
        g = op.gate
        m = Measurement(g.key, tuple(op.qubits), g.full_invert_mask())
        axes = [self.index[q] for q in m.qubits]
        k = len(axes)
        probs = engine.marginal_batch(states, self.n, self.dm, R, axes)
        idx, p_sel = engine.draw_rows(probs, u, self.ws)
        bits = ((idx[:, None] >> (k - 1 - np.arange(k))[None, :]) & 1).astype(np.uint8)
        mask = 0
        values = np.zeros(R, dtype=np.uint64)
        for col, a in enumerate(axes):
            b = bits[:, col].astype(np.uint64)
            if self.dm:
                rb, cb = 2 * self.n - 1 - a, self.n - 1 - a
                mask |= (1 << rb) | (1 << cb)
                values |= (b << np.uint64(rb)) | (b << np.uint64(cb))
            else:
                bb = self.n - 1 - a
                mask |= 1 << bb
                values |= b << np.uint64(bb)
        p64 = p_sel.astype(np.float64)
        with np.errstate(divide="ignore"):
            scales = (1.0 / p64) if self.dm else (1.0 / np.sqrt(p64))
        engine.collapse_batch(states, self.n_bits, R, mask, values, scales.astype(np.float32))
        inv = np.zeros(k, dtype=np.uint8)
        for i, x in enumerate(m.invert_mask[:k]):
            inv[i] = 1 if x else 0
        return m.key, bits ^ inv[None, :]

    def _kraus_chain(self, states, R: int, chain: list, u_dev: torch.Tensor):
        """Kraus channels in a row: each one's weights, the walk and the chosen operator, all on the
        device.  For one-qubit channels after the first, applying the previous channel's choice and
        computing this channel's weights share one sweep of the copies."""
        prev = None  # (targets, device matrices, count, selections, scales) of the last channel
        for j in chain:
            targets, m_dev, n_mats, _ = self._channels[j]
            if prev is None:
                W = engine.kraus_weights_batch_dev(states, self.n_bits, R, targets, m_dev, n_mats, self.ws,
                                                   to_host=False)
            else:
                pt, pm, pn, ps, pc = prev
                W = engine.select_then_kraus_weights(states, self.n_bits, R, pt[0], pm, pn, ps, pc, targets[0],
                                                     m_dev, n_mats, self.ws)
            sel, scl = engine.kraus_choose_batch(W, u_dev[j])
            prev = (targets, m_dev, n_mats, sel, scl)
        targets, m_dev, n_mats, sel, scl = prev
        engine.apply_select_batch_dev(states, self.n_bits, R, targets, m_dev, n_mats, sel, scl)


__all__ = ["TrajectoryBatch", "choice_rows", "kraus_draws"]
