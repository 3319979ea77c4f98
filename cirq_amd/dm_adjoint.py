"""Gradients of Tr(H rho) for noisy density-matrix circuits: adjoint sweep with sqrt checkpointing.

The density matrix is the 2n-bit vector the density-matrix engine evolves (row bits above column
bits, as cirq-core/cirq/sim/density_matrix_simulation_state.py:177-178 puts the left axes i and the
right axes i+n).  Every lowered op S is linear in rho: a gate is U on the row bits and conj(U) on
the column bits, a channel is sum_k K (x) conj(K) on the column and row bits
(cirq-core/cirq/protocols/apply_channel_protocol.py:297-356).  With lambda = vec(H) walked
backwards (lambda <- S^dagger lambda),

    d Tr(H rho) / dt = Re sum_k <lambda_k | dS_k/dt | rho_{k-1}>,

summed over the row and column ops of every gate occurrence that depends on t.  Gates are undone
by their inverses as in the state-vector adjoint sweep.  Channels are not invertible in general,
so every channel that cannot be undone ends a step (the unitary ops and undoable channels up to
the next such channel); the forward pass keeps a copy of rho every ceil(sqrt(steps)) steps and the
reverse sweep recomputes each stretch of steps from its copy, keeping rho at the end of every
step of that stretch.  Memory is about 2 sqrt(steps) density matrices; work is two forward
simulations and one reverse sweep.

A channel whose superoperator S is invertible and well conditioned (weak depolarizing or
amplitude-damping noise: singular values 1 - 4p/3, 1 - gamma) is undone in place instead: rho
takes S^-1 and lambda takes S^dagger, which commute with every gradient op on other bits just as a
gate's U^dagger pair does.  Rounding errors in rho grow by at most ||S^-1|| per undone channel, so
channels are undone, walking backwards, only while the product of those norms since the last
restored copy stays within `max_amplification`; the rest end steps.  invert_channels=False
checkpoints at every channel.

Cirq has no gradients: its users difference simulate_expectation_values_sweep calls
(cirq-core/cirq/sim/simulator.py:349-371) on DensityMatrixSimulator.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib, engine
from ._lib import check, lib, ptr
from .adjoint import _dagger, _grad_op, _inverse, adjoint_waves, default_tile_bits
from .gates import is_parameterized
from .paulis import PauliSum, adjoint_scale, pauli_masks
from .qubits import QubitOrder
from .study import ParamResolver
from .tape import (
    LOp,
    Lowering,
    Pass,
    PackedProgram,
    ParamEvaluator,
    ParamRef,
    SlotChain,
    fill_param_mats,
    fuse_blocks,
    fuse_single_bit,
    order_tile_bits,
    plan_passes,
    schedule,
)


def _inverse_norm(op: LOp) -> float:
    """||S^-1|| (spectral norm) of a channel or scale op, inf when S is singular."""
    if op.kind == "scale":
        c = abs(complex(op.mat))
        return 1.0 / c if c > 0 else math.inf
    smin = float(np.linalg.svd(np.asarray(op.mat, dtype=np.complex128), compute_uv=False)[-1])
    return 1.0 / smin if smin > 1e-12 else math.inf


def _undo_pair(op: LOp) -> list[LOp]:
    """S^-1 on rho and S^dagger on lambda for an invertible channel or scale op (same bits, so
    the scheduler keeps them in one pass)."""
    if op.kind == "scale":
        c = complex(op.mat)
        return [LOp("scale", (), np.complex128(1 / c), (), None, _lib.SEL_PSI),
                LOp("scale", (), np.complex128(np.conj(c)), (), None, _lib.SEL_PHI)]
    m = np.asarray(op.mat, dtype=np.complex128)
    return [LOp(op.kind, op.targets, np.linalg.inv(m), op.controls, None, _lib.SEL_PSI),
            LOp(op.kind, op.targets, _dagger(op.kind, m), op.controls, None, _lib.SEL_PHI)]


def _adjoint_of_nonunitary(op: LOp) -> LOp:
    """S^dagger of a channel or scale op, applied to lambda only.  It is a scheduling fence: moved
    ahead of a gradient op of its step (even one on other bits) it would change that op's
    <lambda|dS|rho>, since unlike a gate it is not applied to rho as well."""
    if op.kind == "scale":
        return LOp("scale", (), np.conj(op.mat), (), None, _lib.SEL_PHI, fence=True)
    return LOp(op.kind, op.targets, _dagger(op.kind, op.mat), op.controls, None, _lib.SEL_PHI, fence=True)


class DMAdjointProgram:
    """Host side of a density-matrix adjoint gradient; needs no GPU.

    steps[j] = (channel that starts the step or None, [(lowered op index, op), ...] up to the next
    such channel); `undone` holds the indices of the channels undone in place.  The forward
    program and the reversed program are each one packed program whose passes are grouped by
    step: fwd_range[j] / rev_range[j] are the pass ranges of step j.  Reversed step j undoes the
    step's ops on rho (psi) and lambda (phi) with gradient ops in between, then applies the
    adjoint of the channel that starts the step to lambda only.
    """

    def __init__(self, circuit, observable, qubit_order=QubitOrder.DEFAULT, *, tile_bits: int | None = None,
                 adjoint_tile_bits: int | None = None, coalesce_bits: int = 4, mfma_blocks: bool = True,
                 checkpoint_every: int | None = None, invert_channels: bool = True,
                 max_amplification: float = 2.0):
        qubits = QubitOrder.as_qubit_order(qubit_order).order_for(circuit.all_qubits())
        self.qubits = tuple(qubits)
        n = len(qubits)
        self.n = n
        nb = 2 * n
        self.n_bits = nb
        low = Lowering(qubits, dm=True, split_dm_diag=True)
        low.lower_circuit(circuit)
        if low.measurements:
            raise ValueError("density-matrix gradients need a circuit without measurements")
        qmap = {q: i for i, q in enumerate(qubits)}
        psum = PauliSum.wrap(observable)
        for t in psum:
            for q in t.qubits:
                if q not in qmap:
                    raise KeyError(q)
        self.terms = [pauli_masks(t, qmap, n) for t in psum]
        # H / s on the row bits of vec(I) gives lambda = vec(H) / s (gradients are multiplied by s)
        self.phi_scale = adjoint_scale(c for _, _, c in self.terms)
        self.row_terms = [(int(x) << n, int(z) << n, c / self.phi_scale) for x, z, c in self.terms]
        self.evaluator = ParamEvaluator(low.param_refs)
        self.symbols = list(self.evaluator.symbols)

        # walking backwards, undo channels while the accumulated ||S^-1|| allows; the others
        # end steps (rho is restored from a stored copy there, which resets the growth)
        cut: set = set()
        self.undone: set = set()
        growth = 1.0
        for i in reversed(low.nonunitary):
            a = _inverse_norm(low.ops[i]) if invert_channels else math.inf
            if growth * a <= max_amplification:
                growth *= a
                self.undone.add(i)
            else:
                cut.add(i)
                growth = 1.0
        steps: list[tuple[LOp | None, list[tuple[int, LOp]]]] = [(None, [])]
        for i, op in enumerate(low.ops):
            if i in cut:
                steps.append((op, []))
            else:
                steps[-1][1].append((i, op))
        self.steps = steps
        S = len(steps)
        self.checkpoint_every = max(1, int(checkpoint_every) if checkpoint_every else math.isqrt(S - 1) + 1)

        dL, dLa = default_tile_bits(nb)
        L = min(tile_bits if tile_bits is not None else dL, nb)
        La = min(adjoint_tile_bits if adjoint_tile_bits is not None else dLa, nb)
        fwd_passes: list[Pass] = []
        rev_passes: list[Pass] = []
        self.fwd_range: list[tuple[int, int]] = []
        self.rev_range: list[tuple[int, int]] = []
        self.grad_slots: list[tuple[ParamRef, int]] = []
        for chan, region in steps:
            ops = ([chan] if chan is not None else []) + [op for _, op in region]
            a = len(fwd_passes)
            if ops:
                fwd_passes += plan_passes(ops, nb, L=L, coalesce=coalesce_bits, blocks=mfma_blocks)
            self.fwd_range.append((a, len(fwd_passes)))
            rev: list[LOp] = []
            for i, op in reversed(region):
                if i in self.undone:
                    rev += _undo_pair(op)
                    continue
                if op.fixed:
                    rev.append(_inverse(op, _lib.SEL_BOTH))
                    continue
                rev.append(_inverse(op, _lib.SEL_PSI))
                for k, e in enumerate(op.pref.exprs):
                    if is_parameterized(e):
                        slot = len(self.grad_slots)
                        self.grad_slots.append((op.pref, k))
                        rev.append(_grad_op(op, k, slot))
                rev.append(_inverse(op, _lib.SEL_PHI))
            if chan is not None:
                rev.append(_adjoint_of_nonunitary(chan))
            a = len(rev_passes)
            if rev:
                ps = schedule(fuse_single_bit(rev), nb, La, coalesce_bits)
                if mfma_blocks:
                    ps = [Pass(p.L, p.lpos, fuse_blocks(p.ops, set(p.lpos), p.L)) for p in ps]
                rev_passes += [order_tile_bits(p, coalesce_bits) for p in ps]
            self.rev_range.append((a, len(rev_passes)))
        self.fwd = PackedProgram(fwd_passes, nb)
        self.rev = PackedProgram(rev_passes, nb)
        self.partial_stride = (1 << (nb - La)) * adjoint_waves(La)
        self._chain = SlotChain([ref.exprs[k] for ref, k in self.grad_slots], self.symbols, self.evaluator)

    @property
    def n_slots(self) -> int:
        return len(self.grad_slots)

    @property
    def n_steps(self) -> int:
        return len(self.steps)

    def segments(self) -> list[tuple[int, int]]:
        """Step ranges [a, b) that start at a checkpoint, in forward order."""
        s, S = self.checkpoint_every, len(self.steps)
        return [(a, min(a + s, S)) for a in range(0, S, s)]

    def symbol_values(self, values) -> np.ndarray:
        if isinstance(values, (ParamResolver, dict)):
            return self.evaluator.symbol_values([ParamResolver(values)])
        v = np.asarray(values, dtype=np.float64).reshape(-1)
        if v.size != len(self.symbols):
            raise ValueError(f"expected {len(self.symbols)} symbol values, got {v.size}")
        return v.reshape(-1, 1).copy()

    def mats(self, symvals: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        return (fill_param_mats(self.fwd, self.evaluator, symvals)[0],
                fill_param_mats(self.rev, self.evaluator, symvals)[0])

    def chain(self, gs: np.ndarray, symvals: np.ndarray) -> np.ndarray:
        """d Tr(H rho)/d symbol from the per-slot sums 2 Re <lambda|dS|rho> with lambda = vec(H) / s
        (each slot is one row or column op, so the gradient is s/2 times their chained sum)."""
        gs = np.asarray(gs, dtype=np.float64).reshape(1, -1) * (0.5 * self.phi_scale)
        return self._chain(gs, symvals)[0]


class DMAdjointGradient:
    """Tr(H rho) and its gradient for a parameterised noisy circuit on one GPU.

    circuit: sympy-parameterised gates and channels, no measurements; observable: PauliSumLike.
    """

    def __init__(self, circuit, observable, qubit_order=QubitOrder.DEFAULT, **kw):
        self.device = engine.require_gpu()
        self.prog = DMAdjointProgram(circuit, observable, qubit_order, **kw)
        self.symbols = self.prog.symbols
        self.n = self.prog.n
        nb = self.prog.n_bits
        self.dev_fwd = engine.DeviceProgram(self.prog.fwd, self.device)
        self.dev_rev = engine.DeviceProgram(self.prog.rev, self.device, adjoint=True)
        self.ws = engine.Workspace(self.device)
        self.rho = engine.new_state(nb, self.device)
        self.lam = torch.empty_like(self.rho)
        segs = self.prog.segments()
        self.checkpoints = [torch.empty_like(self.rho) for _ in range(len(segs) - 1)]
        self.stretch = [torch.empty_like(self.rho) for _ in range(max(b - a for a, b in segs) - 1)]
        G = self.prog.n_slots
        self.partial = torch.zeros(max(G, 1) * self.prog.partial_stride, dtype=torch.float64, device=self.device)
        self.gslot = torch.zeros(max(G, 1), dtype=torch.float64, device=self.device)
        self.fwd_spec = self.rev_spec = None  # specialised kernels (specialize())

    def specialize(self):
        """Compiles the forward and reverse programs to straight-line kernels (cirq_amd.specialize)
        and runs them from now on."""
        from .specialize import specialize

        self.fwd_spec = specialize(self.prog.fwd)
        self.rev_spec = specialize(self.prog.rev, adjoint=True)
        return self

    @staticmethod
    def _handles(spec, a: int):
        """Launch handles of passes a, a+1, ... of a specialised program."""
        return ctypes.c_void_p(ctypes.addressof(spec.handles) + a * ctypes.sizeof(ctypes.c_void_p))

    def _init_rho(self, buf: torch.Tensor):
        check(lib().cqs_init_basis(ptr(buf), self.prog.n_bits, 0, engine.stream_handle()), "cqs_init_basis")

    def _run_step(self, j: int, buf: torch.Tensor):
        a, b = self.prog.fwd_range[j]
        if b <= a:
            return
        d = self.dev_fwd
        args = (ptr(buf), self.prog.n_bits)
        rest = (d.pass_arr[a:b].ctypes.data, b - a, ptr(d.ops), ptr(d.mats), ptr(d.aux))
        if self.fwd_spec is not None:
            check(lib().cqs_run_passes_specialized(*args, -1, *rest, self._handles(self.fwd_spec, a),
                                                   engine.stream_handle()), "cqs_run_passes_specialized")
        else:
            check(lib().cqs_run_passes(*args, *rest, engine.stream_handle()), "cqs_run_passes")

    def _reverse_step(self, j: int, psi: torch.Tensor):
        a, b = self.prog.rev_range[j]
        if b <= a:
            return
        d = self.dev_rev
        args = (ptr(psi), ptr(self.lam), self.prog.n_bits, d.pass_arr[a:b].ctypes.data, b - a, ptr(d.ops), ptr(d.mats),
                ptr(d.aux), ptr(self.partial), self.prog.partial_stride)
        if self.rev_spec is not None:
            check(lib().cqs_adjoint_passes_specialized(*args, self._handles(self.rev_spec, a), engine.stream_handle()),
                  "cqs_adjoint_passes_specialized")
        else:
            check(lib().cqs_adjoint_passes(*args, engine.stream_handle()), "cqs_adjoint_passes")

    def _set_mats(self, values) -> np.ndarray:
        symvals = self.prog.symbol_values(values)
        mf, mr = self.prog.mats(symvals)
        self.dev_fwd.set_mats(mf)
        self.dev_rev.set_mats(mr)
        return symvals

    def forward_rho(self, values) -> torch.Tensor:
        """rho(values) as a 2n-bit vector (a view of this object's buffer)."""
        self._set_mats(values)
        self._init_rho(self.rho)
        for j in range(self.prog.n_steps):
            self._run_step(j, self.rho)
        return self.rho

    def value(self, values) -> float:
        """Tr(H rho) alone: the forward passes and one expectation, no reverse sweep."""
        rho = self.forward_rho(values)
        return float(engine.expect_terms(rho, self.n, True, self.prog.terms, self.ws).sum().real.item())

    def value_and_grad(self, values) -> tuple[float, np.ndarray]:
        """(Tr(H rho), d Tr(H rho)/d symbol in `symbols` order) at the given symbol values."""
        prog, n = self.prog, self.n
        symvals = self._set_mats(values)
        segs = prog.segments()
        # forward, keeping rho at the start of every segment but the first
        self._init_rho(self.rho)
        for g, (a, b) in enumerate(segs):
            if g > 0:
                self.checkpoints[g - 1].copy_(self.rho)
            for j in range(a, b):
                self._run_step(j, self.rho)
        value = float(engine.expect_terms(self.rho, n, True, prog.terms, self.ws).sum().real.item())
        # lambda = vec(H): H on the row bits of vec(I)
        eye = self.rho
        eye.zero_()
        eye.view(1 << n, 1 << n).diagonal().fill_(1.0)
        engine.apply_paulisum(eye, self.lam, prog.n_bits, prog.row_terms)
        # reverse: recompute each segment from its checkpoint, then walk its steps backwards
        for g in range(len(segs) - 1, -1, -1):
            a, b = segs[g]
            if g == 0:
                self._init_rho(self.rho)
            else:
                self.rho.copy_(self.checkpoints[g - 1])
            for j in range(a, b - 1):
                self._run_step(j, self.rho)
                self.stretch[j - a].copy_(self.rho)
            self._run_step(b - 1, self.rho)
            self._reverse_step(b - 1, self.rho)
            for j in range(b - 2, a - 1, -1):
                self._reverse_step(j, self.stretch[j - a])
        G = prog.n_slots
        check(lib().cqs_grad_reduce(ptr(self.partial), G, prog.partial_stride, ptr(self.gslot), engine.stream_handle()),
              "cqs_grad_reduce")
        gs = self.gslot.cpu().numpy()[:G]
        return value, prog.chain(gs, symvals)
