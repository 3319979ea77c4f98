"""Batched sweeps: expectation values (and adjoint gradients) of many parameterised circuits at
many parameter points, all states resident on the GPU at once.

Cirq evaluates a sweep one point at a time (`simulate_sweep_iter`,
cirq-core/cirq/sim/simulator_base.py:275-320; `simulate_expectation_values_sweep_iter`,
cirq-core/cirq/sim/sparse_simulator.py:193-218): resolve, copy the initial state, simulate,
reduce.  Here every (circuit, resolver) point owns a state in one arena, and each pass level of
every program runs as one launch over all points that have it (`cqs_batch_passes`), so a
4096-point sweep over circuits of 10-22 qubits costs a few hundred launches instead of a few
hundred per point.  The launch sequence is captured once into a HIP graph and replayed; a new set
of parameter values only rewrites the matrix arena.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib, engine
from ._lib import check, lib, ptr
from .adjoint import AdjointProgram
from .paulis import PauliSum, adjoint_scale, pauli_masks
from .tape import fill_param_mats

CHUNK_BITS = 12  # amplitudes per workgroup of the init / expectation / H psi kernels: 2^12


def _align(x: int, a: int) -> int:
    return -(-x // a) * a


def _batch_pass_rows(packed, op_base: int, aux_base: int) -> np.ndarray:
    n = packed.n_bits
    rows = np.zeros(packed.n_passes, dtype=_lib.BATCH_PASS_DTYPE)
    for i, ps in enumerate(packed.pass_arr):
        L = int(ps["L"])
        r = rows[i]
        r["n_bits"], r["L"], r["nouter"] = n, L, n - L
        r["op_begin"] = op_base + int(ps["op_begin"])
        r["op_count"] = int(ps["op_count"])
        r["max_k"] = int(ps["max_k"])
        r["lpos"][:L] = ps["lpos"][:L]
        r["lpos_out"][:L] = ps["lpos_out"][:L]
        loc = set(int(x) for x in ps["lpos"][:L])
        outer = [p for p in range(n) if p not in loc]
        r["opos"][: len(outer)] = outer
    return rows


def _rebased_ops(packed, aux_base: int) -> np.ndarray:
    ops = packed.ops_arr[: packed.n_ops].copy()
    kind = ops["kind"] & 0xFF
    uses_aux = (kind == _lib.OP_PAIRPHASE) | (kind == _lib.OP_PERMUTE)
    ops["t"][uses_aux, 0] += aux_base
    return ops


class BatchPrograms:
    """Host side of a batch: compiled programs, the shared op / aux / pass tables, and (after
    `plan`) every descriptor array of a run over `counts[c]` points of circuit c.  Needs no GPU."""

    def __init__(self, circuits, observables, *, gradients: bool = True, tile_bits: int | None = None,
                 adjoint_tile_bits: int | None = None, coalesce_bits: int = 4, mfma_blocks: bool = True):
        if len(circuits) != len(observables):
            raise ValueError("one observable list per circuit")
        self.gradients = gradients
        self.progs: list[AdjointProgram] = []
        self.term_masks: list[list] = []
        for c, obs in zip(circuits, observables):
            terms = [t for o in obs for t in PauliSum.wrap(o)]
            prog = AdjointProgram(c, PauliSum.from_pauli_strings(terms) if terms else PauliSum(),
                                  tile_bits=tile_bits, adjoint_tile_bits=adjoint_tile_bits,
                                  coalesce_bits=coalesce_bits, mfma_blocks=mfma_blocks)
            qmap = {q: i for i, q in enumerate(prog.qubits)}
            self.term_masks.append([pauli_masks(t, qmap, prog.n) for t in terms])
            self.progs.append(prog)
        self.symbols = [p.symbols for p in self.progs]
        ops, aux, rows = [], [], []
        self.fwd_pass0, self.rev_pass0 = [], []
        n_ops = n_aux = n_rows = 0
        for prog in self.progs:
            for packed, first in ((prog.fwd, self.fwd_pass0), (prog.rev, self.rev_pass0)):
                first.append(n_rows)
                ops.append(_rebased_ops(packed, n_aux))
                rows.append(_batch_pass_rows(packed, n_ops, n_aux))
                aux.append(packed.aux_arr)
                n_ops += packed.n_ops
                n_aux += packed.aux_arr.size
                n_rows += packed.n_passes
        self.ops = np.concatenate(ops)
        self.aux = np.concatenate(aux).astype(np.int32)
        self.passes = np.concatenate(rows)
        xm, zm, cf, acf, tbase = [], [], [], [], []
        self.phi_scale = []
        for masks in self.term_masks:
            tbase.append(len(xm))
            s = adjoint_scale(co for _, _, co in masks)
            self.phi_scale.append(s)
            for x, z, co in masks:
                xm.append(x)
                zm.append(z)
                cf.append(complex(co))
                acf.append(complex(co) / s)
        self.term_base = tbase
        self.xmask = np.array(xm or [0], dtype=np.uint64)
        self.zmask = np.array(zm or [0], dtype=np.uint64)
        self.coef = np.array(cf or [0j], dtype=np.complex128)
        # phi = (H_c / s_c) psi for the adjoint sweeps; the slot gradients are multiplied by s_c
        self.apply_coef = np.array(acf or [0j], dtype=np.complex128)
        self.counts = None

    def plan(self, counts: tuple):
        progs = self.progs
        pts = [(c, j) for c in range(len(progs)) for j in range(counts[c])]
        self.points = pts
        st_off, st2_off, off = [], [], 0
        for c, _ in pts:
            size = 1 << progs[c].n
            st_off.append(off)
            off += size
            st2_off.append(off if self.gradients else 0)
            off += size if self.gradients else 0
        self.arena_elems = max(off, 1)
        mf_off, mr_off, moff = [], [], 0
        for c, _ in pts:
            mf_off.append(moff)
            moff += _align(progs[c].fwd.mats_arr.size, 32)
            mr_off.append(moff)
            moff += _align(progs[c].rev.mats_arr.size, 32) if self.gradients else 0
        self.mats_elems = max(moff, 1)
        self.mf_off, self.mr_off = mf_off, mr_off
        part_off, poff = [], 0
        for c, _ in pts:
            part_off.append(poff)
            poff += progs[c].n_slots * progs[c].partial_stride if self.gradients else 0
        self.partial_elems = max(poff, 1)
        launches, items_all = [], []
        n_items_so_far = [0]

        def push(kind, L, arr, n_tiles):
            launches.append((kind, L, n_items_so_far[0], len(arr), n_tiles))
            items_all.append(arr)
            n_items_so_far[0] += len(arr)

        init = np.zeros(len(pts), dtype=_lib.BATCH_ITEM_DTYPE)
        tb = 0
        for i, (c, _) in enumerate(pts):
            init[i]["state"] = st_off[i]
            init[i]["pass"] = self.fwd_pass0[c]
            init[i]["tile_begin"] = tb
            n = progs[c].n
            tb += 1 << (n - min(CHUNK_BITS, n))
        push("init", 0, init, tb)

        def add_levels(kind, packed_of, first_row, mats_off):
            depth = max(packed_of(progs[c]).n_passes for c, _ in pts) if pts else 0
            for lev in range(depth):
                by_L: dict = {}
                for i, (c, _) in enumerate(pts):
                    pk = packed_of(progs[c])
                    if lev < pk.n_passes:
                        by_L.setdefault(int(pk.pass_arr["L"][lev]), []).append(i)
                for L, members in sorted(by_L.items()):
                    arr = np.zeros(len(members), dtype=_lib.BATCH_ITEM_DTYPE)
                    tb = 0
                    for k, i in enumerate(members):
                        c = pts[i][0]
                        arr[k]["state"] = st_off[i]
                        arr[k]["state2"] = st2_off[i]
                        arr[k]["mats"] = mats_off[i]
                        arr[k]["partial"] = part_off[i]
                        arr[k]["pass"] = first_row[c] + lev
                        arr[k]["tile_begin"] = tb
                        tb += 1 << (progs[c].n - L)
                    check(lib().cqs_batch_validate(
                        self.passes.ctypes.data, len(self.passes), len(self.ops), arr.ctypes.data, len(arr), L,
                        self.arena_elems, self.mats_elems, int(kind == "adj")), f"batch {kind} level {lev}")
                    push(kind, L, arr, tb)

        add_levels("fwd", lambda p: p.fwd, self.fwd_pass0, mf_off)
        if self.gradients:
            add_levels("adj", lambda p: p.rev, self.rev_pass0, mr_off)
        self.launches = launches
        self.items = np.concatenate(items_all)
        titems = np.zeros(len(pts), dtype=_lib.BATCH_TERMS_DTYPE)
        ev_seg = [0]
        bb = out = 0
        for i, (c, _) in enumerate(pts):
            n = progs[c].n
            nb = 1 << (n - min(CHUNK_BITS, n))
            nt = len(self.term_masks[c])
            titems[i] = (st_off[i], st2_off[i], out, n, self.term_base[c], nt, bb)
            for _t in range(nt):
                ev_seg.append(ev_seg[-1] + nb)
            out += nt * nb
            bb += nb
        self.term_items = titems
        self.n_term_blocks = bb
        self.ev_partial_elems = max(out, 1)
        self.ev_seg = np.array(ev_seg, dtype=np.int64)
        g_seg = [0]
        if self.gradients:
            for c, _ in pts:
                for _s in range(progs[c].n_slots):
                    g_seg.append(g_seg[-1] + progs[c].partial_stride)
        self.g_seg = np.array(g_seg, dtype=np.int64)
        self.counts = counts

    def fill_mats(self, symvals: list, out: np.ndarray):
        """Every point's forward (and reverse) matrix buffer into `out` (complex64 arena)."""
        i = 0
        for c, prog in enumerate(self.progs):
            B = self.counts[c]
            if B == 0:
                continue
            mf = fill_param_mats(prog.fwd, prog.evaluator, symvals[c])
            mr = fill_param_mats(prog.rev, prog.evaluator, symvals[c]) if self.gradients else None
            for b in range(B):
                out[self.mf_off[i] : self.mf_off[i] + mf.shape[1]] = mf[b]
                if mr is not None:
                    out[self.mr_off[i] : self.mr_off[i] + mr.shape[1]] = mr[b]
                i += 1

    def results(self, raw_ev: np.ndarray, slot_grads: np.ndarray | None, symvals: list) -> tuple[list, list]:
        evs, grads = [], []
        t = gpos = 0
        for c, prog in enumerate(self.progs):
            B, nt = self.counts[c], len(self.term_masks[c])
            co = self.coef[self.term_base[c] : self.term_base[c] + nt]
            evs.append((raw_ev[t : t + B * nt].reshape(B, nt) * co[None, :]).real)
            t += B * nt
            if self.gradients:
                slots = slot_grads[gpos : gpos + B * prog.n_slots].reshape(B, prog.n_slots) * self.phi_scale[c]
                gpos += B * prog.n_slots
                grads.append(prog.chain_batch(slots, symvals[c]))
            else:
                grads.append(np.zeros((B, 0)))
        return evs, grads


class BatchedExpectation:
    """<P_t> for every Pauli term t of a circuit's observable list, and optionally the adjoint
    gradient of sum_t <P_t> (the terms carry their weights as coefficients), for many points.

    circuits: list of parameterised unitary circuits; observables: per circuit, a list of Pauli
    strings (PauliSumLike terms).  `run(values)` takes, per circuit, an array (points, symbols)
    of symbol values in `symbols[c]` order.
    """

    def __init__(self, circuits, observables, *, gradients: bool = True, use_graph: bool = True, **kw):
        self.device = engine.require_gpu()
        self.host = BatchPrograms(circuits, observables, gradients=gradients, **kw)
        self.gradients = gradients
        self.use_graph = use_graph
        self.symbols = self.host.symbols
        h = self.host
        self.ops = torch.from_numpy(h.ops.view(np.int32).copy()).to(self.device)
        self.aux = torch.from_numpy(h.aux.copy()).to(self.device)
        self.passes = torch.from_numpy(h.passes.view(np.uint8).copy()).to(self.device)
        self.xm = torch.from_numpy(h.xmask.view(np.int64).copy()).to(self.device)
        self.zm = torch.from_numpy(h.zmask.view(np.int64).copy()).to(self.device)
        self.cf = torch.from_numpy(h.apply_coef.view(np.float64).copy()).to(self.device)
        self._graph = None

    def _plan(self, counts: tuple):
        h = self.host
        h.plan(counts)
        dev = self.device
        self.arena = torch.empty(h.arena_elems, dtype=torch.complex64, device=dev)
        self.mats_host = torch.zeros(h.mats_elems, dtype=torch.complex64).pin_memory()
        self.mats = torch.zeros(h.mats_elems, dtype=torch.complex64, device=dev)
        self.partial = torch.zeros(h.partial_elems, dtype=torch.float64, device=dev)
        self.items = torch.from_numpy(h.items.view(np.uint8).copy()).to(dev)
        self.term_items = torch.from_numpy(h.term_items.view(np.uint8).copy()).to(dev)
        self.ev_partial = torch.zeros(2 * h.ev_partial_elems, dtype=torch.float64, device=dev)
        self.ev_seg = torch.from_numpy(h.ev_seg).to(dev)
        self.ev_out = torch.zeros(2 * max(len(h.ev_seg) - 1, 1), dtype=torch.float64, device=dev)
        self.g_seg = torch.from_numpy(h.g_seg).to(dev)
        self.g_out = torch.zeros(max(len(h.g_seg) - 1, 1), dtype=torch.float64, device=dev)
        self._graph = None

    def _expect(self, s):
        h = self.host
        check(lib().cqs_batch_expect(ptr(self.arena), ptr(self.term_items), len(h.term_items), h.n_term_blocks,
                                     ptr(self.xm), ptr(self.zm), ptr(self.ev_partial), CHUNK_BITS, s),
              "cqs_batch_expect")
        check(lib().cqs_segment_sum(ptr(self.ev_partial), ptr(self.ev_seg), len(h.ev_seg) - 1, 2, ptr(self.ev_out),
                                    s), "cqs_segment_sum")

    def _launch_all(self):
        """init, forward levels, <P_t>, phi = H psi, reverse levels, gradient reduction."""
        h = self.host
        s = engine.stream_handle()
        L_ = lib()
        item_bytes = _lib.BATCH_ITEM_DTYPE.itemsize
        base_items = ptr(self.items)
        ops, aux, passes, mats, arena = ptr(self.ops), ptr(self.aux), ptr(self.passes), ptr(self.mats), ptr(self.arena)
        adj_started = False
        for kind, L, off, n_items, n_tiles in h.launches:
            items = base_items + off * item_bytes
            if kind == "init":
                check(L_.cqs_batch_init(arena, items, n_items, n_tiles, passes, CHUNK_BITS, s), "cqs_batch_init")
            elif kind == "fwd":
                check(L_.cqs_batch_passes(arena, items, n_items, n_tiles, L, passes, ops, mats, aux, s),
                      "cqs_batch_passes")
            else:
                if not adj_started:
                    adj_started = True
                    self._expect(s)
                    self._apply(s)
                check(L_.cqs_batch_adjoint(arena, items, n_items, n_tiles, L, passes, ops, mats, aux,
                                           ptr(self.partial), s), "cqs_batch_adjoint")
        if not adj_started:
            self._expect(s)
        if self.gradients:
            check(L_.cqs_segment_sum(ptr(self.partial), ptr(self.g_seg), len(h.g_seg) - 1, 1, ptr(self.g_out), s),
                  "cqs_segment_sum")

    def _apply(self, s):
        h = self.host
        check(lib().cqs_batch_apply_paulisum(ptr(self.arena), ptr(self.term_items), len(h.term_items),
                                             h.n_term_blocks, ptr(self.xm), ptr(self.zm), ptr(self.cf), CHUNK_BITS,
                                             s), "cqs_batch_apply_paulisum")

    def run(self, values) -> tuple[list, list]:
        """values[c]: (points_c, n_symbols_c) array.  Returns (evs, grads): per circuit, arrays of
        shape (points_c, n_terms_c) (real expectation values, weights included) and
        (points_c, n_symbols_c) (d sum_t <P_t> / d symbol; empty without gradients)."""
        h = self.host
        vals = [np.asarray(v, dtype=np.float64).reshape(-1, len(self.symbols[c])) for c, v in enumerate(values)]
        counts = tuple(v.shape[0] for v in vals)
        if counts != h.counts:
            self._plan(counts)
        symvals = [v.T.copy() for v in vals]  # (n_symbols, points)
        h.fill_mats(symvals, self.mats_host.numpy())
        self.mats.copy_(self.mats_host, non_blocking=True)
        if self.use_graph:
            if self._graph is None:
                self._launch_all()  # warm-up outside capture (LDS attributes, lazy loads)
                torch.cuda.synchronize()
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g):
                    self._launch_all()
                self._graph = g
            self._graph.replay()
        else:
            self._launch_all()
        raw = self.ev_out.cpu().numpy().view(np.complex128)
        gs = self.g_out.cpu().numpy() if self.gradients else None
        return h.results(raw, gs, symvals)
