"""Gradients of expectation values: one adjoint sweep on the GPU, plus a parameter-shift baseline.

Cirq has no gradient machinery: a Cirq user differentiates <psi(theta)|H|psi(theta)> by calling
simulate_expectation_values_sweep (cirq-core/cirq/sim/simulator.py:349-371) at shifted parameter
values.  AdjointGradient instead returns the value and d<H>/d(symbol) for every symbol from one
forward simulation and one reverse sweep (Jones & Gacon 2020, arXiv:2009.02823): the forward state
psi and phi = H psi are walked back gate by gate, un-computing each gate with its inverse rather than
storing intermediate states, and every parameterised gate contributes 2 Re <phi|dU/dt|psi>.

ExpectationLayer wraps it in a torch.autograd.Function.
"""

from __future__ import annotations

import ctypes

import numpy as np
import sympy
import torch

from . import _lib, engine
from ._lib import check, lib, ptr
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
    compile_ops,
    embed,
    fill_param_mats,
    fuse_blocks,
    fuse_single_bit,
    order_tile_bits,
    schedule,
)


def _dagger(kind: str, x):
    """Inverse payload of an op: conjugate transpose (dense) or conjugate (diagonal / phases)."""
    if x is None:
        return None
    if kind in ("dense", "block"):
        return np.conj(np.swapaxes(x, -1, -2))
    return np.conj(x)


def _inverse(op: LOp, sel: int) -> LOp:
    """U^dagger on the states `sel` selects.  The one-state inverses of a parameterised op and its
    GRAD ops must share a pass: GRAD ops fold the controls into their matrix and need them local,
    so the inverses ask for the same local bits.  (Split across passes, another op's gradient
    could run while only one of the two states had this inverse applied.)"""
    if op.kind == "scale" and abs(abs(complex(op.mat)) - 1) > 1e-9:
        raise ValueError("adjoint gradients need a unitary circuit (non-unit scale factor)")
    ref = op.pref
    if ref is not None:
        base, kind = ref.transform, op.kind
        ref = ParamRef(ref.gate, lambda b, base=base, kind=kind: _dagger(kind, base(b)), ref.exprs, ref.slot)
    need = tuple(b for b, _ in op.controls) if sel != _lib.SEL_BOTH else ()
    return LOp(op.kind, op.targets, _dagger(op.kind, op.mat), op.controls, ref, sel, need=need)


def _grad_op(op: LOp, k: int, slot: int) -> LOp:
    """d(op)/d(parameter k) as a GRAD op on the op's bits.

    The lowering's payload transforms are affine in the gate matrix (a controlled diagonal puts
    constant ones outside the controlled block), so the derivative payload is
    transform(dM) - transform(0).  Controls of a dense op are folded into its matrix the same way:
    the blocks where the controls fail do not depend on the parameter."""
    ref = op.pref
    base = ref.transform

    def linear(b, base=base):
        return base(b) - base(np.zeros_like(b))

    if op.kind == "diag":
        return LOp("grad_diag", op.targets, None, (), ParamRef(ref.gate, linear, ref.exprs, ref.slot, k), 0, slot)
    if op.kind != "dense":
        raise TypeError(f"cannot differentiate a {op.kind} op")
    if not op.controls:
        return LOp("grad_dense", op.targets, None, (), ParamRef(ref.gate, linear, ref.exprs, ref.slot, k), 0, slot)
    tb = list(op.targets)
    ctrls = tuple(op.controls)
    bits = tb + [b for b, _ in ctrls]

    def controlled(b, linear=linear, tb=tb, ctrls=ctrls, bits=bits):
        d = linear(b)
        zero = embed(np.zeros_like(d[0]), tb, ctrls, bits)  # identity where the controls fail
        return np.stack([embed(x, tb, ctrls, bits) - zero for x in d])

    return LOp("grad_dense", tuple(bits), None, (), ParamRef(ref.gate, controlled, ref.exprs, ref.slot, k), 0, slot)


def default_tile_bits(n_bits: int) -> tuple[int, int]:
    """(forward, reverse) tile widths of a gradient over n_bits state bits.  Small states have few
    tiles, one workgroup or less per CU at 12-bit tiles, so narrower tiles buy occupancy at the
    price of more passes; the reverse kernel holds two states and runs best one bit narrower.
    Timed on MI355X over ansatz gradients of 14-24 qubits (value_and_grad, ms): 18 qubits
    1.83 (10/10) vs 2.96 (12/12); 20 qubits 2.50 (11/11) vs 3.38; 22 qubits 7.73 (12/11) vs 7.86;
    24 qubits 20.7 (12/11) vs 21.8."""
    if n_bits <= 18:
        return min(10, n_bits), min(10, n_bits)
    if n_bits <= 21:
        return 11, 11
    return 12, 11


def adjoint_waves(L: int) -> int:
    """Waves per workgroup of the adjoint kernels at tile width L: each writes its own partial of
    every GRAD op (256 threads up to L = 12, then 16 elements per thread)."""
    return (256 if L <= 12 else (1 << L) // 16) // 64


def shift_rule_ok(gate, k: int) -> bool:
    """True when parameter k of `gate` enters as exp(i pi t h) with two eigenvalues h one apart,
    the case where the two-term shift rule with shifts +-1/2 is exact."""
    from .gates import EigenGate, PhasedXPowGate

    if isinstance(gate, EigenGate):
        hs = sorted({round(float(h), 12) for h, _ in gate._eigen_components()})
        return len(hs) == 2 and abs(hs[1] - hs[0] - 1) < 1e-12
    return isinstance(gate, PhasedXPowGate) and k == 1


class AdjointProgram:
    """Host side of an adjoint gradient: the forward program, the reversed program and the map
    from gradient slots to parameter expressions.  Needs no GPU.

    The reversed program walks the lowered ops backwards.  A fixed op U becomes U^dagger on both
    states; a parameterised op becomes [U^dagger on psi][dU/dt_k for each symbolic parameter
    t_k][U^dagger on phi].  Ops on disjoint bits commute, so the scheduler may move whole ops
    across each other; each triple keeps its order because its members share their bits.
    """

    def __init__(self, circuit, observable, qubit_order=QubitOrder.DEFAULT, *, tile_bits: int | None = None,
                 adjoint_tile_bits: int | None = None, coalesce_bits: int = 4, mfma_blocks: bool = True):
        qubits = QubitOrder.as_qubit_order(qubit_order).order_for(circuit.all_qubits())
        self.qubits = tuple(qubits)
        n = len(qubits)
        self.n = n
        low = Lowering(qubits, dm=False)
        low.lower_circuit(circuit)
        if low.measurements:
            raise ValueError("adjoint gradients need a circuit without measurements")
        qmap = {q: i for i, q in enumerate(qubits)}
        psum = PauliSum.wrap(observable)
        for t in psum:
            for q in t.qubits:
                if q not in qmap:
                    raise KeyError(q)
        self.terms = [pauli_masks(t, qmap, n) for t in psum]
        # phi = (H' / s) psi with H' = H minus its identity terms: c I adds c d<psi|psi>/dt = 0
        # to every derivative in exact arithmetic and only rounding noise in complex64
        varying = [(x, z, c) for x, z, c in self.terms if x or z]
        self.phi_scale = adjoint_scale(c for _, _, c in varying)
        self.phi_terms = [(x, z, c / self.phi_scale) for x, z, c in varying]
        # <H> = s Re<psi|phi> + c_I <psi|psi> (c_I: the identity terms' coefficient, real for a
        # Hermitian H): two inner products instead of one reduction per group of terms
        self.identity_coef = float(sum(complex(c).real for x, z, c in self.terms if not x and not z))
        self.evaluator = ParamEvaluator(low.param_refs)
        self.symbols = list(self.evaluator.symbols)
        dL, dLa = default_tile_bits(n)
        L = min(tile_bits if tile_bits is not None else dL, n)
        La = min(adjoint_tile_bits if adjoint_tile_bits is not None else dLa, n)
        self.fwd = compile_ops(low.ops, n, L=L, coalesce=coalesce_bits, blocks=mfma_blocks)
        rev: list[LOp] = []
        self.grad_slots: list[tuple[ParamRef, int]] = []  # slot -> (forward op's ref, parameter index)
        for op in reversed(low.ops):
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
        rev = fuse_single_bit(rev)
        passes = schedule(rev, n, La, coalesce_bits)
        if mfma_blocks:
            passes = [Pass(p.L, p.lpos, fuse_blocks(p.ops, set(p.lpos), p.L)) for p in passes]
        passes = [order_tile_bits(p, coalesce_bits) for p in passes]
        self.rev = PackedProgram(passes, n)
        self.tile_bits_rev = La
        self.partial_stride = (1 << (n - La)) * adjoint_waves(La)  # one FP64 partial per wave of each tile
        self._slot_exprs = [ref.exprs[k] for ref, k in self.grad_slots]
        self._chain = SlotChain(self._slot_exprs, self.symbols, self.evaluator)

    @property
    def n_slots(self) -> int:
        return len(self.grad_slots)

    def symbol_values(self, values) -> np.ndarray:
        """(n_symbols, 1) array from a resolver/dict or a vector in `symbols` order."""
        if isinstance(values, (ParamResolver, dict)):
            return self.evaluator.symbol_values([ParamResolver(values)])
        v = np.asarray(values, dtype=np.float64).reshape(-1)
        if v.size != len(self.symbols):
            raise ValueError(f"expected {len(self.symbols)} symbol values, got {v.size}")
        return v.reshape(-1, 1).copy()

    def mats(self, symvals: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """Matrix buffers of the forward and reversed programs at the given symbol values."""
        return (fill_param_mats(self.fwd, self.evaluator, symvals)[0],
                fill_param_mats(self.rev, self.evaluator, symvals)[0])

    def chain(self, gs: np.ndarray, symvals: np.ndarray) -> np.ndarray:
        """d<H>/d symbol from per-slot derivatives d<H>/d t_slot (one point)."""
        return self.chain_batch(np.asarray(gs)[None, :], symvals)[0]

    def chain_batch(self, gs: np.ndarray, symvals: np.ndarray) -> np.ndarray:
        """(B, n_symbols) gradients from (B, n_slots) slot derivatives; symvals (n_symbols, B)."""
        return self._chain(gs, symvals)

    def shifted_mats(self, symvals: np.ndarray) -> np.ndarray:
        """(2 * n_slots, mats) forward matrix buffers with parameter `slot` of one gate occurrence
        shifted by +1/2 (even rows) and -1/2 (odd rows) in exponent units."""
        base = fill_param_mats(self.fwd, self.evaluator, symvals)[0]
        out = np.broadcast_to(base, (2 * self.n_slots, base.size)).copy()
        entries: dict = {}
        for off, r, size in self.fwd.param_entries:
            entries.setdefault(r.slot, []).append((off, r, size))
        for slot, (ref, k) in enumerate(self.grad_slots):
            if not shift_rule_ok(ref.gate, k):
                raise ValueError(f"the two-term shift rule does not hold for parameter {k} of {ref.gate!r}")
            vals = [self.evaluator.values(e, symvals) for e in ref.exprs]
            vv = [np.repeat(v, 2) for v in vals]
            vv[k] = vv[k] + np.array([0.5, -0.5])
            mb = ref.gate._matrix_batch(tuple(vv))
            for off, r, size in entries.get(ref.slot, ()):  # every lowered op of this gate occurrence
                out[2 * slot : 2 * slot + 2, off : off + size] = r.transform(mb).reshape(2, -1)
        return out


class AdjointGradient:
    """<H> and d<H>/d(symbol) of a parameterised unitary circuit on one GPU.

    circuit: sympy-parameterised gates, no measurements; observable: PauliSumLike.
    """

    def __init__(self, circuit, observable, qubit_order=QubitOrder.DEFAULT, **kw):
        self.device = engine.require_gpu()
        self.prog = AdjointProgram(circuit, observable, qubit_order, **kw)
        self.symbols = self.prog.symbols
        self.n = self.prog.n
        self.dev_fwd = engine.DeviceProgram(self.prog.fwd, self.device)
        self.dev_rev = engine.DeviceProgram(self.prog.rev, self.device, adjoint=True)
        self.ws = engine.Workspace(self.device)
        self.psi = engine.new_state(self.n, self.device)
        self.phi = torch.empty_like(self.psi)
        G = self.prog.n_slots
        self.partial = torch.zeros(max(G, 1) * self.prog.partial_stride, dtype=torch.float64, device=self.device)
        self.gslot = torch.zeros(max(G, 1), dtype=torch.float64, device=self.device)
        self.ev_buf = torch.zeros(4, dtype=torch.float64, device=self.device)
        self.rev_spec = None  # specialised reverse-sweep kernels (specialize())

    def specialize(self):
        """Compiles the forward and reverse programs to straight-line kernels (cirq_amd.specialize)
        and runs them from now on."""
        from .specialize import specialize

        self.dev_fwd.specialize()
        self.rev_spec = specialize(self.prog.rev, adjoint=True)
        return self

    def _forward(self):
        check(lib().cqs_init_basis(ptr(self.psi), self.n, 0, engine.stream_handle()), "cqs_init_basis")
        self.dev_fwd.run(self.psi)

    def forward_state(self, values) -> torch.Tensor:
        """The forward state psi(values) (a view of this object's buffer)."""
        symvals = self.prog.symbol_values(values)
        self.dev_fwd.set_mats(fill_param_mats(self.prog.fwd, self.prog.evaluator, symvals)[0])
        self._forward()
        return self.psi

    def value(self, values) -> float:
        """<H> alone: the forward passes and the expectation, no reverse sweep (the same arithmetic
        as value_and_grad's value)."""
        self.forward_state(values)
        return self._expectation_value(self._queue_expectation().cpu().numpy())

    def _queue_expectation(self, out: torch.Tensor | None = None) -> torch.Tensor:
        """Queues phi = (H' / s) psi and the inner products <psi|phi>, <psi|psi> into a device
        float64[4] (`out`, default this object's buffer); _expectation_value turns them into
        <H> = s Re<psi|phi> + c_I <psi|psi>."""
        prog, n = self.prog, self.n
        out = self.ev_buf if out is None else out
        engine.apply_paulisum(self.psi, self.phi, n, prog.phi_terms)
        engine.inner(self.psi, self.phi, n, self.ws, out=out[0:2])
        if prog.identity_coef:
            engine.inner(self.psi, self.psi, n, self.ws, out=out[2:4])
        return out

    def _expectation_value(self, buf: np.ndarray) -> float:
        ev = float(buf[0]) * self.prog.phi_scale
        if self.prog.identity_coef:
            ev += self.prog.identity_coef * float(buf[2])
        return ev

    def value_and_grad(self, values, timings: dict | None = None) -> tuple[float, np.ndarray]:
        """(<H>, d<H>/d symbol in `symbols` order) at the given symbol values.  With a `timings`
        dict, also fills timings['forward_ms'] / ['reverse_ms']: the forward and reverse programs'
        GPU time between events on the stream they run on."""
        prog, n = self.prog, self.n
        symvals = prog.symbol_values(values)
        mf, mr = prog.mats(symvals)
        self.dev_fwd.set_mats(mf)
        self.dev_rev.set_mats(mr)
        ev_t = [torch.cuda.Event(enable_timing=True) for _ in range(4)] if timings is not None else None
        if ev_t:
            ev_t[0].record()
        self._forward()
        if ev_t:
            ev_t[1].record()
        evbuf = self._queue_expectation()
        if ev_t:
            ev_t[2].record()
        G = prog.n_slots
        rv = prog.rev
        args = (ptr(self.psi), ptr(self.phi), n, rv.pass_arr.ctypes.data, int(rv.n_passes), ptr(self.dev_rev.ops),
                ptr(self.dev_rev.mats), ptr(self.dev_rev.aux), ptr(self.partial), prog.partial_stride)
        if self.rev_spec is not None:
            check(lib().cqs_adjoint_passes_specialized(
                *args, ctypes.cast(self.rev_spec.handles, ctypes.c_void_p), engine.stream_handle()),
                "cqs_adjoint_passes_specialized")
        else:
            check(lib().cqs_adjoint_passes(*args, engine.stream_handle()), "cqs_adjoint_passes")
        if ev_t:
            ev_t[3].record()
        check(lib().cqs_grad_reduce(ptr(self.partial), G, prog.partial_stride, ptr(self.gslot), engine.stream_handle()),
              "cqs_grad_reduce")
        ev = self._expectation_value(evbuf.cpu().numpy())
        if ev_t:
            timings["forward_ms"] = ev_t[0].elapsed_time(ev_t[1])
            timings["reverse_ms"] = ev_t[2].elapsed_time(ev_t[3])
        gs = self.gslot.cpu().numpy()[:G] * prog.phi_scale
        return ev, prog.chain(gs, symvals)

    def parameter_shift(self, values) -> tuple[float, np.ndarray]:
        """The same gradient by the two-term shift rule: two forward simulations and <H> per
        symbolic parameter of every gate occurrence, all matrix buffers resident on the device
        and the expectation values read back once."""
        prog, n = self.prog, self.n
        symvals = prog.symbol_values(values)
        shifted = prog.shifted_mats(symvals)
        base = prog.mats(symvals)[0]
        rows = np.vstack([shifted, base[None]])
        width = -(-rows.shape[1] // 32) * 32  # keeps every row's MFMA operands 16-byte aligned
        padded = np.zeros((rows.shape[0], width), dtype=np.complex64)
        padded[:, : rows.shape[1]] = rows
        allm = torch.from_numpy(padded.view(np.float32)).to(self.device)
        own = self.dev_fwd.mats
        bufs = torch.zeros((allm.shape[0], 4), dtype=torch.float64, device=self.device)
        try:
            for r in range(allm.shape[0]):  # <H> as value_and_grad computes it
                self.dev_fwd.set_mats_device(allm[r])
                self._forward()
                self._queue_expectation(bufs[r])
            ev = np.array([self._expectation_value(b) for b in bufs.cpu().numpy()])
        finally:
            self.dev_fwd.set_mats_device(own)
        gs = (np.pi / 2) * (ev[0:-1:2] - ev[1:-1:2])
        return float(ev[-1]), prog.chain(gs, symvals)


class _ExpectationFn(torch.autograd.Function):
    """Rows of symbol values (B, S) -> <H> per row (B,); the adjoint sweep runs in forward and its
    gradients are kept for backward."""

    @staticmethod
    def forward(ctx, values: torch.Tensor, grad_engine):
        rows = values.detach().cpu().numpy().reshape(-1, values.shape[-1])
        vg = [grad_engine.value_and_grad(r) for r in rows]
        grads = np.stack([g for _, g in vg])
        ctx.save_for_backward(torch.as_tensor(grads, dtype=values.dtype, device=values.device).reshape(values.shape))
        return torch.tensor([v for v, _ in vg], dtype=values.dtype, device=values.device).reshape(values.shape[:-1])

    @staticmethod
    def backward(ctx, gout):
        (grad,) = ctx.saved_tensors
        return gout.unsqueeze(-1) * grad, None


class ExpectationLayer(torch.nn.Module):
    """torch module: symbol values (in `.symbols` order; shape (S,) or a batch (B, S)) -> <H>
    (a scalar, or (B,)), differentiable.  When no gradient is needed (torch.no_grad, or inputs
    that do not require grad) only the forward passes run; otherwise one adjoint sweep per row.

    density_matrix=True evolves rho instead (the circuit may hold channels) and differentiates
    Tr(H rho) with the checkpointed density-matrix sweep of cirq_amd.dm_adjoint."""

    def __init__(self, circuit, observable, qubit_order=QubitOrder.DEFAULT, *, density_matrix: bool = False, **kw):
        super().__init__()
        if density_matrix:
            from .dm_adjoint import DMAdjointGradient

            self.engine = DMAdjointGradient(circuit, observable, qubit_order, **kw)
        else:
            self.engine = AdjointGradient(circuit, observable, qubit_order, **kw)
        self.symbols = self.engine.symbols

    def forward(self, values: torch.Tensor) -> torch.Tensor:
        if values.dim() not in (1, 2) or values.shape[-1] != len(self.symbols):
            raise ValueError(f"expected (S,) or (B, S) symbol values with S = {len(self.symbols)}, got {tuple(values.shape)}")
        if torch.is_grad_enabled() and values.requires_grad:
            return _ExpectationFn.apply(values, self.engine)
        rows = values.detach().cpu().numpy().reshape(-1, values.shape[-1])
        out = [self.engine.value(r) for r in rows]
        return torch.tensor(out, dtype=values.dtype, device=values.device).reshape(values.shape[:-1])

