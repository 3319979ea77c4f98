#!/usr/bin/env python3
"""Benchmark of the state-vector hot path on one MI355X.

Workload (one step): the 30-qubit random grid circuit on a 5x6 grid, 40 cycles of
{sqrt(X), sqrt(Y), sqrt(W)} single-qubit layers and CZ layers A,B,C,D (1,200 single-qubit gates +
490 CZ), complex64, started from |0...0>, then 10^6 samples of all qubits (bit-exact NumPy-choice
sampling in Cirq's factor order) and <Z_i Z_j> for the 49 nearest-neighbour pairs.

Metric: layers (cycles) per second of the whole step; HBM GB/s of the gate-pass kernel.
Prints one JSON line.  --gpus N with N > 1 exits: the simulator runs on one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def _chain_bytes(packed, n_bits: int) -> float:
    """Algorithmic HBM bytes of one run of a program from a basis state.  A pass reads only the
    tiles the state can reach by then (those whose bits that no earlier pass had local equal the
    basis state's; the first pass builds the state in registers and reads nothing), and it writes
    the tiles the next pass reads plus the zeros of the rest, except where the next pass writes
    those zeros itself: it writes what the next pass reads, and the last pass writes everything."""
    full = 8.0 * 2**n_bits
    fixed = (1 << n_bits) - 1
    zmasks = []
    for ps in packed.pass_arr[: int(packed.n_passes)]:
        if int(ps["op_count"]) <= 0:
            continue
        local = 0
        for b in range(int(ps["L"])):
            local |= 1 << int(ps["lpos"][b])
        zmasks.append(fixed & ~local)
        fixed &= ~local
    total = 0.0
    for i, z in enumerate(zmasks):
        read = 0.0 if i == 0 else full / 2 ** bin(z).count("1")
        write = full / 2 ** bin(zmasks[i + 1]).count("1") if i + 1 < len(zmasks) else full
        total += read + write
    return total


def _progress(msg: str) -> None:
    """One line per finished leg on stderr (stdout carries only the result line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--rows", type=int, default=5)
    ap.add_argument("--cols", type=int, default=6)
    ap.add_argument("--depth", type=int, default=40)
    ap.add_argument("--reps", type=int, default=1_000_000)
    ap.add_argument("--tile-bits", type=int, default=None)
    ap.add_argument("--coalesce-bits", type=int, default=None)
    ap.add_argument("--no-blocks", action="store_true", help="disable MFMA block fusion")
    ap.add_argument("--no-autotune", action="store_true",
                    help="compile the main circuit with the default tile / fusion settings instead of timing candidates")
    ap.add_argument("--no-specialize", action="store_true",
                    help="run the grid legs' passes through the op interpreter instead of compiled per-pass kernels")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-qubits", type=int, default=26)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint-gradient leg")
    ap.add_argument("--adjoint-qubits", type=int, default=20)
    ap.add_argument("--adjoint-layers", type=int, default=10)
    ap.add_argument("--adjoint-reps", type=int, default=5)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-sized device buffer 5 times (reference for HBM byte counters) and exit")
    ap.add_argument("--no-qaoa", action="store_true", help="skip the QAOA gradient leg")
    ap.add_argument("--qaoa-nodes", type=int, default=24)
    ap.add_argument("--qaoa-p", type=int, default=8)
    ap.add_argument("--no-dm", action="store_true", help="skip the noisy density-matrix leg")
    ap.add_argument("--dm-qubits", type=int, default=16)
    ap.add_argument("--dm-layers", type=int, default=20)
    ap.add_argument("--dm-cpu-qubits", type=int, default=12)
    ap.add_argument("--no-q32", action="store_true", help="skip the 32-qubit depth-20 grid leg")
    ap.add_argument("--q32-steps", type=int, default=2)
    ap.add_argument("--no-dm-adjoint", action="store_true", help="skip the noisy density-matrix gradient leg")
    ap.add_argument("--dm-adjoint-qubits", type=int, default=12)
    ap.add_argument("--dm-adjoint-layers", type=int, default=10)
    ap.add_argument("--no-trajectories", action="store_true", help="skip the noisy-trajectory leg")
    ap.add_argument("--traj-qubits", type=int, default=16)
    ap.add_argument("--traj-layers", type=int, default=10)
    ap.add_argument("--traj-reps", type=int, default=1024)
    ap.add_argument("--traj-looped-reps", type=int, default=3)
    ap.add_argument("--no-sweep", action="store_true", help="skip the batched-sweep leg")
    ap.add_argument("--sweep-circuits", type=int, default=256)
    ap.add_argument("--sweep-points", type=int, default=16)
    ap.add_argument("--sweep-looped-circuits", type=int, default=256,
                    help="circuits whose first point the looped comparison runs (scaled to all points)")
    ap.add_argument("--leg", choices=sorted(LEGS), default=None,
                    help="run only this leg (no main workload) and print its record: for profiling one leg")
    return ap.parse_args()


def dm_leg(args):
    """Noisy density matrix: 16-qubit line (2^32 complex64 entries, 32 GiB), 20 layers of Rz, Ry per
    qubit and CZ on alternating pairs, depolarize(1e-3) after every gate, amplitude_damp(1e-3) on
    every qubit each layer; then Tr(rho sum_i Z_i)."""
    import torch

    import cirq_amd as cirq
    from cirq_amd import engine, generators
    from cirq_amd.paulis import pauli_masks

    n = args.dm_qubits
    circuit = generators.noisy_line_circuit(n, args.dm_layers, seed=0)
    qubits = sorted(circuit.all_qubits())
    sim = cirq.DensityMatrixSimulator(specialize_passes=not args.no_specialize)
    t0 = time.perf_counter()
    comp = sim._compile(circuit, qubits)
    t_compile = time.perf_counter() - t0
    _progress(f"density-matrix program compiled in {t_compile:.1f} s")
    qmap = {q: i for i, q in enumerate(qubits)}
    terms = [pauli_masks(t, qmap, n) for t in generators.sum_z(qubits)]
    rho = engine.new_state(2 * n, sim._dev())
    stream = torch.cuda.current_stream()

    def once():
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        comp.dev.run(rho, basis=0)  # |0..0><0..0|, built by the first pass
        e1.record(stream)
        ev = engine.expect_terms(rho, n, True, terms, sim._ws).sum().real.item()
        return e0.elapsed_time(e1), ev

    once()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ms, ev = once()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    n_pass = int(comp.packed.n_passes)
    bytes_per_pass = _chain_bytes(comp.packed, 2 * n) / max(n_pass, 1)  # from |0..0><0..0|
    del rho
    torch.cuda.empty_cache()
    out = {
        "workload": f"noisy line {n} qubits x {args.dm_layers} layers, density matrix 2^{2 * n} complex64, "
        f"{sum(1 for _ in circuit.all_operations())} operations incl. channels",
        "layers_per_s": args.dm_layers / wall,
        "passes": n_pass,
        "passes_ms": ms,
        "pass_avg_ms": ms / max(n_pass, 1),
        "pass_hbm_gbs": bytes_per_pass / (ms / 1e3 / max(n_pass, 1)) / 1e9,
        "specialized_passes": comp.dev.specialized is not None,
        "compile_s": t_compile,
        "tr_rho_sum_z": ev,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = dm_cpu_baseline(args.dm_cpu_qubits)
    return out


def dm_cpu_baseline(n: int) -> dict:
    """The oracle's density-matrix step (apply_channel with left / right Kraus products, complex64,
    one NumPy thread) at n qubits: each op kind of the noisy-line layer (Rz, Ry, CZ, depolarize,
    amplitude damping) timed once on the first layer's ops (median of 3 after a warm-up); a
    layer's time is their sum weighted by the layer's op counts."""
    from oracle import cirq_np as O
    from tests.helpers import to_oracle_ops
    from cirq_amd import generators

    circuit = generators.noisy_line_circuit(n, 1, seed=0)
    ops = to_oracle_ops(circuit, sorted(circuit.all_qubits()))
    rho = np.zeros((2**n, 2**n), dtype=np.complex64)
    rho[0, 0] = 1
    state = {"t": rho.reshape((2,) * (2 * n))}
    kinds, counts = {}, {}
    for op in ops:
        kinds.setdefault(op[1], op)
        counts[op[1]] = counts.get(op[1], 0) + 1

    def apply(op):
        kind, fam, params, axes = op
        ks = [O.gate_unitary(fam, params)] if kind == "gate" else O.channel_kraus(fam, params)
        state["t"] = O.apply_channel(state["t"], [k.astype(np.complex64) for k in ks], list(axes), n)

    per = {fam: _median_s(lambda op=op: apply(op), 3) for fam, op in kinds.items()}
    layer = sum(per[f] * counts[f] for f in per)
    return {
        "value": 1.0 / layer, "unit": "layers/s", "cores": 1, "kind": "port",
        "sample": f"noisy-line layer at {n} qubits (2^{2 * n}-entry density matrix) through the oracle's "
        f"apply_channel: per-op medians {', '.join(f'{f} {t * 1e3:.0f} ms x{counts[f]}' for f, t in per.items())}; "
        f"{layer:.2f} s per layer ({_cpu_threads()})",
    }


def trajectory_leg(args):
    """run() of a noisy circuit on a state vector: every channel is sampled per repetition, as
    Cirq's simulator does.  Batched (all repetitions as copies of one state) against one
    repetition at a time (a sample of repetitions, scaled), each after a warm-up call that
    compiles its programs."""
    import cirq_amd as cirq
    from cirq_amd import generators

    circuit = generators.noisy_line_circuit(args.traj_qubits, args.traj_layers, seed=0)
    qs = sorted(circuit.all_qubits())
    from cirq_amd.tape import _has_channel

    n_channels = sum(1 for op in circuit.all_operations() if _has_channel(op))
    circuit = circuit + cirq.Circuit(cirq.measure(*qs, key="m"))
    sim = cirq.Simulator(seed=1)
    t0 = time.perf_counter()
    sim.run(circuit, repetitions=args.traj_reps)
    first = time.perf_counter() - t0
    t0 = time.perf_counter()
    res = sim.run(circuit, repetitions=args.traj_reps)
    batched = time.perf_counter() - t0
    one = cirq.Simulator(seed=1, trajectory_batch=1)
    one.run(circuit, repetitions=1)
    t0 = time.perf_counter()
    one.run(circuit, repetitions=args.traj_looped_reps)
    looped = (time.perf_counter() - t0) / args.traj_looped_reps
    ones = res.records["m"][:, 0, :].mean()
    return {
        "workload": f"noisy line {args.traj_qubits} qubits x {args.traj_layers} layers on a state vector, "
        f"{n_channels} sampled channels per repetition, {args.traj_reps} repetitions",
        "batched_s": batched,
        "batched_reps_per_s": args.traj_reps / batched,
        "first_call_s": first,
        "looped_reps_per_s": 1.0 / looped,
        "looped_sample": f"{args.traj_looped_reps} repetitions one at a time ({looped * args.traj_looped_reps:.2f} s)",
        "speedup": looped * args.traj_reps / batched,
        "mean_measured_bit": float(ones),
    }


def sweep_leg(args):
    """Mixed-size batched sweep: 256 hardware-efficient ansatz circuits (10-22 qubits, 4-16 layers),
    16 points each (4096), per point <P_t> of 32 random 3-local strings with N(0,1) weights plus
    the adjoint gradient of their sum.  Batched (one launch per pass level over all points, HIP
    graph replay) vs looped (one point at a time through the single-state path)."""
    import torch

    from cirq_amd import generators
    from cirq_amd.adjoint import AdjointGradient
    from cirq_amd.batch import BatchedExpectation
    from cirq_amd.paulis import PauliSum

    circuits, observables, values = generators.mixed_sweep_workload(
        n_circuits=args.sweep_circuits, points=args.sweep_points, seed=0)
    n_points = sum(v.shape[0] for v in values)
    t0 = time.perf_counter()
    be = BatchedExpectation(circuits, observables)
    t_compile = time.perf_counter() - t0
    _progress(f"sweep programs compiled in {t_compile:.1f} s")
    be.run(values)  # plan, graph capture, warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    evs, grads = be.run(values)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    _progress(f"batched sweep {t_batch:.2f} s")
    # looped: the first point of each of the first K circuits, one at a time
    K = min(args.sweep_looped_circuits, len(circuits))
    ags = [AdjointGradient(circuits[c], PauliSum.from_pauli_strings(observables[c])) for c in range(K)]
    for c in range(K):  # every looped circuit warmed (first launches, workspace growth) before timing
        ags[c].value_and_grad(values[c][1 % values[c].shape[0]])
    torch.cuda.synchronize()
    diffs = []
    t0 = time.perf_counter()
    for c in range(K):
        v, g = ags[c].value_and_grad(values[c][0])
        diffs.append(max(float(np.abs(g - grads[c][0]).max()), abs(v - float(evs[c][0].sum()))))
    torch.cuda.synchronize()
    t_loop_sample = time.perf_counter() - t0
    looped_points = K
    t_loop = t_loop_sample / looped_points * n_points
    qubits = [len(c.all_qubits()) for c in circuits]
    return {
        "workload": f"{len(circuits)} HEA circuits x {args.sweep_points} points = {n_points} points, "
        f"{min(qubits)}-{max(qubits)} qubits, 32 random 3-local strings each, <P_t> + adjoint gradients",
        "batched_s": t_batch,
        "batched_points_per_s": n_points / t_batch,
        "looped_points_per_s": looped_points / t_loop_sample,
        "looped_sample": f"first point of {looped_points} circuits ({t_loop_sample:.2f} s), scaled to {n_points}",
        "speedup": t_loop / t_batch,
        "max_abs_diff_batched_vs_looped": max(diffs),
        "compile_s": t_compile,
        "launches": len(be.host.launches),
        "roofline": _sweep_bytes_roofline(
            [(p.n, int(p.fwd.n_passes), int(p.rev.n_passes), int(v.shape[0])) for p, v in zip(be.host.progs, values)],
            t_batch, "every point's forward and reverse passes over the batched wall time (one replay of the "
            "captured launches); states of 10-22 qubits stay in the last-level cache"),
    }


def _specialize_gradient(args, ag):
    """Compiles the gradient's forward and reverse programs to straight-line kernels (timed; both
    the adjoint sweep and the parameter-shift comparison then run them), unless --no-specialize."""
    if args.no_specialize:
        return None
    t0 = time.perf_counter()
    ag.specialize()
    return time.perf_counter() - t0


def _adjoint_roofline(ag, n_bits: int, point) -> dict:
    """Roofline of the reverse-sweep kernel (k_adjoint_pass, the dominant one of a gradient): its
    algorithmic bytes per pass (psi and phi each read and written once: 4 * 8 * 2^n) over its
    average pass time, timed with HIP events on the stream the passes run on; the forward passes
    beside it (read + write once each, the first writes only)."""
    t = {}
    ag.value_and_grad(point, timings=t)
    n_rev, n_fwd = int(ag.prog.rev.n_passes), int(ag.prog.fwd.n_passes)
    rev_bytes = 4 * 8 * 2**n_bits
    fwd_bytes = (2 * n_fwd - 1) * 8 * 2**n_bits / max(n_fwd, 1)
    achieved = rev_bytes / (t["reverse_ms"] / 1e3 / max(n_rev, 1)) / 1e9
    fwd_gbs = fwd_bytes / (t["forward_ms"] / 1e3 / max(n_fwd, 1)) / 1e9
    state_mib = 8 * 2**n_bits / 2**20
    where = ("the two states stay in the 256 MiB last-level cache" if 2 * state_mib <= 128 else
             "the two states fill the 256 MiB last-level cache and spill to HBM")
    return {
        "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
        "traffic": None, "kernel": f"k_adjoint_pass<{int(ag.prog.tile_bits_rev)}>",
        "reverse_ms": t["reverse_ms"], "reverse_passes": n_rev, "forward_ms": t["forward_ms"],
        "forward_gbs": fwd_gbs,
        "note": f"{state_mib:.0f} MiB per state, {2**(n_bits - int(ag.prog.tile_bits_rev))} tiles per reverse pass: "
        f"{where}; the passes are bound by VALU work per op (about 3 interpreted ops per parameterised gate) "
        "and by latency at this tile count, not by HBM bandwidth",
    }


def adjoint_leg(args):
    """Hardware-efficient ansatz (Ry, Rz per qubit + CZ ladder per layer; 20 qubits x 10 layers =
    400 symbols), H = sum_i Z_i + 32 random 4-local strings: gradients/s of one adjoint sweep vs
    the GPU parameter-shift rule (2 forward simulations + <H> per symbol), and their max
    difference."""
    import torch

    from cirq_amd import generators
    from cirq_amd.adjoint import AdjointGradient

    circuit, syms, vals = generators.hardware_efficient_ansatz(args.adjoint_qubits, args.adjoint_layers, seed=0)
    psum = generators.random_pauli_sum(sorted(circuit.all_qubits()), 32, 4, seed=0)
    ag = AdjointGradient(circuit, psum)
    t_spec = _specialize_gradient(args, ag)
    ag.value_and_grad(vals)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.adjoint_reps):
        v, g = ag.value_and_grad(vals)
    torch.cuda.synchronize()
    t_adj = (time.perf_counter() - t0) / args.adjoint_reps
    ag.parameter_shift(vals)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    v2, g2 = ag.parameter_shift(vals)
    torch.cuda.synchronize()
    t_ps = time.perf_counter() - t0
    G = len(syms)
    out = {
        "workload": f"HEA {args.adjoint_qubits} qubits x {args.adjoint_layers} layers, {G} symbols, "
        f"{len(list(psum))} Pauli terms",
        "adjoint_ms": t_adj * 1e3,
        "adjoint_grads_per_s": G / t_adj,
        "param_shift_ms": t_ps * 1e3,
        "param_shift_grads_per_s": G / t_ps,
        "speedup": t_ps / t_adj,
        "max_abs_grad_diff": float(np.abs(g - g2).max()),
        "value_diff": abs(v - v2),
        "forward_passes": int(ag.prog.fwd.n_passes),
        "reverse_passes": int(ag.prog.rev.n_passes),
        "roofline": _adjoint_roofline(ag, args.adjoint_qubits, vals),
        "specialize_s": t_spec,
    }
    if not args.no_cpu_baseline:
        # Cirq has no adjoint: its gradient is the shift rule, 2 simulations per symbol
        n_layers = args.adjoint_layers
        layer = _oracle_layer_seconds(circuit, syms, vals, moments=slice(0, len(circuit) // n_layers))
        t_cpu = 2 * G * n_layers * layer
        out["cpu_baseline"] = {
            "value": G / t_cpu, "unit": "gradients/s", "cores": 1, "kind": "port",
            "sample": f"one ansatz layer at {args.adjoint_qubits} qubits through the oracle (median of 3 after a "
            f"warm-up): {layer:.3f} s; the shift rule's {2 * G} simulations of {n_layers} layers scaled from it "
            f"(expectation values not counted; {_cpu_threads()})",
        }
    return out


def _oracle_layer_seconds(circuit, syms, vals, moments: slice) -> float:
    """Median seconds (3 runs after a warm-up) of the oracle applying the ops of circuit[moments]
    at the given symbol values, complex64, from |0...0>."""
    from oracle import cirq_np as O
    from tests.helpers import to_oracle_ops
    from cirq_amd.study import ParamResolver

    qs = sorted(circuit.all_qubits())
    n = len(qs)
    ops = to_oracle_ops(circuit[moments], qs, ParamResolver(dict(zip(syms, vals))))
    psi = np.zeros(2**n, dtype=np.complex64)
    psi[0] = 1
    state = {"t": psi.reshape((2,) * n)}

    def run(sel=ops):
        t = state["t"]
        for _, fam, params, axes in sel:
            t = O.apply_gate(t, fam, params, axes)
        state["t"] = t

    return _median_s(run, 3, warm=lambda: run(ops[:2]))


def qaoa_leg(args):
    """QAOA MaxCut on a seeded random 3-regular graph of 24 nodes, p = 8 (16 symbols, 504 gates),
    H = sum_edges (1 - Z_u Z_v) / 2: gradients/s of one adjoint sweep vs the GPU parameter-shift
    rule (two simulations per parameter occurrence: 960), and their max difference; the CPU
    baseline is the oracle's central differences (2 simulations per symbol, 32)."""
    import torch

    from cirq_amd import generators
    from cirq_amd.adjoint import AdjointGradient

    circuit, syms, vals, ham, edges = generators.qaoa_maxcut(args.qaoa_nodes, args.qaoa_p, seed=0)
    ag = AdjointGradient(circuit, ham)
    t_spec = _specialize_gradient(args, ag)
    point = dict(zip(syms, vals))
    ag.value_and_grad(point)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.adjoint_reps):
        v, g = ag.value_and_grad(point)
    torch.cuda.synchronize()
    t_adj = (time.perf_counter() - t0) / args.adjoint_reps
    ag.parameter_shift(point)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    v2, g2 = ag.parameter_shift(point)
    torch.cuda.synchronize()
    t_ps = time.perf_counter() - t0
    G = len(syms)
    out = {
        "workload": f"QAOA MaxCut, random 3-regular graph of {args.qaoa_nodes} nodes ({len(edges)} edges), "
        f"p = {args.qaoa_p}, {G} symbols, {sum(1 for _ in circuit.all_operations())} gates",
        "value": v,
        "adjoint_ms": t_adj * 1e3,
        "adjoint_grads_per_s": G / t_adj,
        "param_shift_ms": t_ps * 1e3,
        "param_shift_evaluations": 2 * int(ag.prog.n_slots),
        "speedup": t_ps / t_adj,
        "max_abs_grad_diff_vs_param_shift": float(np.abs(g - g2).max()),
        "forward_passes": int(ag.prog.fwd.n_passes),
        "reverse_passes": int(ag.prog.rev.n_passes),
        "roofline": _adjoint_roofline(ag, args.qaoa_nodes, point),
        "specialize_s": t_spec,
    }
    if not args.no_cpu_baseline:
        per_layer = (len(circuit) - 1) // args.qaoa_p  # moments of one (ZZ layer, Rx layer) round
        layer = _oracle_layer_seconds(circuit, syms, vals, moments=slice(1, 1 + per_layer))
        t_sim = (args.qaoa_p + 1.0 / 3) * layer  # the H layer costs about a third of a round
        t_cpu = 2 * G * t_sim
        out["cpu_baseline"] = {
            "value": G / t_cpu, "unit": "gradients/s", "cores": 1, "kind": "port",
            "sample": f"one QAOA round (36 ZZ + {args.qaoa_nodes} Rx) at {args.qaoa_nodes} qubits through the oracle "
            f"(median of 3 after a warm-up): {layer:.2f} s; central differences' {2 * G} simulations scaled from it "
            f"(expectation values not counted; {_cpu_threads()})",
        }
    return out


def dm_adjoint_leg(args):
    """Noisy-circuit gradients: the density-matrix workload with every Rz / Ry angle a symbol
    (12 qubits x 10 layers = 240 symbols, depolarize(1e-3) after every gate, amplitude_damp(1e-3)
    per qubit per layer), H = sum_i Z_i.  One density-matrix adjoint sweep (weak channels undone in
    place), the same sweep checkpointing at every channel (sqrt spacing), and the parameter-shift
    rule through the density-matrix simulator (2 simulations + Tr(H rho) per symbol; the first 8
    symbols timed and scaled), whose gradients pin the adjoint ones."""
    import torch

    import cirq_amd
    from cirq_amd import generators
    from cirq_amd.dm_adjoint import DMAdjointGradient
    from cirq_amd.study import ParamResolver

    n, layers = args.dm_adjoint_qubits, args.dm_adjoint_layers
    circuit, syms, vals = generators.noisy_line_ansatz(n, layers, seed=0)
    H = generators.sum_z(sorted(circuit.all_qubits()))
    G = len(syms)

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            r = fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps, r

    ga = DMAdjointGradient(circuit, H)
    t_spec = _specialize_gradient(args, ga)
    t_adj, (v, g) = timed(lambda: ga.value_and_grad(vals), 3)
    gc = DMAdjointGradient(circuit, H, invert_channels=False)
    _specialize_gradient(args, gc)
    t_ck, (vc, g_ck) = timed(lambda: gc.value_and_grad(vals), 1)
    K = min(8, G)
    base = dict(zip(syms, vals))
    resolvers = []
    for k in range(K):
        for sgn in (1, -1):
            r = dict(base)
            r[syms[k]] = vals[k] + sgn * np.pi / 2
            resolvers.append(ParamResolver(r))
    sim = cirq_amd.DensityMatrixSimulator(specialize_passes=not args.no_specialize)
    t_ps_sample, evs = timed(lambda: sim.simulate_expectation_values_sweep(circuit, H, resolvers), 1)
    ev = np.array([float(np.real(e[0])) for e in evs])
    g_ps = (ev[0::2] - ev[1::2]) / 2
    t_ps = t_ps_sample / (2 * K) * (2 * G)
    return {
        "workload": f"noisy line ansatz {n} qubits x {layers} layers, density matrix 2^{2 * n} complex64, "
        f"{G} symbols, H = sum Z_i",
        "adjoint_ms": t_adj * 1e3,
        "adjoint_grads_per_s": G / t_adj,
        "steps": ga.prog.n_steps,
        "channels_undone": len(ga.prog.undone),
        "forward_passes": int(ga.prog.fwd.n_passes),
        "reverse_passes": int(ga.prog.rev.n_passes),
        "checkpointed_ms": t_ck * 1e3,
        "checkpointed_steps": gc.prog.n_steps,
        "checkpoint_every": gc.prog.checkpoint_every,
        "param_shift_ms": t_ps * 1e3,
        "param_shift_sample": f"first {K} symbols ({t_ps_sample:.3f} s for {2 * K} simulations), scaled to {G}",
        "speedup": t_ps / t_adj,
        "speedup_checkpointed": t_ps / t_ck,
        "max_abs_grad_diff_vs_param_shift": float(np.abs(g[:K] - g_ps).max()),
        "max_abs_grad_diff_checkpointed": float(np.abs(g - g_ck).max()),
        "value_diff_checkpointed": abs(v - vc),
        "specialize_s": t_spec,
        "roofline": _sweep_bytes_roofline(
            [(2 * n, int(ga.prog.fwd.n_passes), int(ga.prog.rev.n_passes), 1)], t_adj,
            "whole gradient: forward passes (rho read + written once) and reverse passes (rho and lambda "
            f"each read + written once) over its wall time; 2^{2 * n} entries = {8 * 4**n / 2**20:.0f} MiB per "
            "matrix, resident in the last-level cache"),
    }


def _sweep_bytes_roofline(items, seconds: float, note: str) -> dict:
    """Whole-job roofline of gradient sweeps: items (n_bits, forward passes, reverse passes,
    points); algorithmic bytes = per point (2 f - 1) * 8 * 2^n (forward; the first pass writes
    only) + 32 * r * 2^n (reverse: two states read and written), over the wall time."""
    total = sum(pts * ((2 * f - 1) * 8 * 2**nb + 32 * r * 2**nb) for nb, f, r, pts in items)
    achieved = total / seconds / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": None, "algorithmic_bytes": total, "note": note}


def _median_s(fn, runs: int = 3, warm=None) -> float:
    """Median wall time of `runs` calls after one warm-up call (`warm`, default `fn`;
    time.perf_counter)."""
    (warm or fn)()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def _cpu_threads() -> str:
    return (f"{len(os.sched_getaffinity(0))} cores visible, OMP_NUM_THREADS={os.environ.get('OMP_NUM_THREADS', 'unset')}, "
            f"OPENBLAS_NUM_THREADS={os.environ.get('OPENBLAS_NUM_THREADS', 'unset')}")


def cpu_baseline(args, rows: int, cols: int, depth: int, n_gpu: int):
    """Cirq's algorithm restated in NumPy (the oracle: each gate's own _apply_unitary_ first, then
    its matrix through slices or einsum), complex64, one thread of NumPy elementwise work, on a
    bounded sample: the first cycle (one single-qubit layer and CZ layer A) of the same circuit
    at args.cpu_qubits qubits, a warm-up (one single-qubit gate and one CZ) then the median of 3
    runs, scaled by the state size to n_gpu qubits."""
    from oracle import cirq_np as O
    from tests.helpers import to_oracle_ops
    from cirq_amd import generators

    n = args.cpu_qubits
    c = generators.grid_circuit(rows, cols, depth, seed=0, n_qubits=n)
    moments = list(c)
    cycle = to_oracle_ops(c[:2], sorted(c.all_qubits()))  # moments 0 and 1: one cycle
    psi = np.zeros(2**n, dtype=np.complex64)
    psi[0] = 1
    state = {"t": psi.reshape((2,) * n)}

    def one_cycle(ops=cycle):
        t = state["t"]
        for _, fam, params, axes in ops:
            t = O.apply_gate(t, fam, params, axes)
        state["t"] = t

    # warm-up: one single-qubit gate and one CZ (touches every page of the state and buffers)
    warm = [cycle[0], next(op for op in cycle if op[1] == "CZPowGate")]
    sec = _median_s(one_cycle, 3, warm=lambda: one_cycle(warm))
    layers_per_s_n = 1.0 / sec
    layers_per_s = layers_per_s_n / (2 ** (n_gpu - n))  # per-op cost scales with the state size
    del state
    return {
        "value": layers_per_s,
        "unit": "layers/s",
        "cores": 1,
        "kind": "port",
        "sample": (
            f"one cycle ({len(cycle)} ops: {sum(1 for op in cycle if op[1] != 'CZPowGate')} single-qubit gates + "
            f"{sum(1 for op in cycle if op[1] == 'CZPowGate')} CZ) of the same circuit at {n} qubits through the oracle "
            f"(Cirq's gate strategies then matrices, complex64, single-threaded NumPy; {_cpu_threads()}): median of 3 "
            f"runs after a warm-up gate pair, {sec:.2f} s per cycle = {layers_per_s_n:.4f} layers/s at {n} qubits, scaled by "
            f"2^-{n_gpu - n} to {n_gpu} qubits"
        ),
        "moments_total": len(moments),
    }


def grid_workload(args, rows: int, cols: int, depth: int, reps: int, steps: int, warmup: int, autotune: bool):
    """One step of the random-grid workload: |0..0>, the circuit's tile passes, `reps` samples of
    every qubit in Cirq's factor order and <Z_i Z_j> over nearest-neighbour pairs.  Times `steps`
    steps after `warmup`, the passes with HIP events on the stream they run on."""
    import torch

    import cirq_amd as cirq
    from cirq_amd import engine, generators
    from cirq_amd.paulis import pauli_masks
    from cirq_amd.simulator import product_state_factors

    n = rows * cols
    circuit = generators.grid_circuit(rows, cols, depth, seed=0)
    qubits = sorted(circuit.all_qubits())
    n_ops = sum(1 for _ in circuit.all_operations())
    zz = generators.nearest_neighbour_zz(qubits)
    qmap = {q: i for i, q in enumerate(qubits)}
    terms = [pauli_masks(t, qmap, n) for z in zz for t in cirq.PauliSum.wrap(z)]

    tuner = None
    if autotune and args.tile_bits is None and args.coalesce_bits is None and not args.no_blocks:
        from cirq_amd.autotune import Autotuner

        tuner = Autotuner(os.environ.get("CIRQ_AMD_TUNE_CACHE"), reps=1)
    # unless --no-specialize, the tuned program's passes run as straight-line kernels
    # (cirq_amd.specialize), planned with the cost entries measured for them
    sim = cirq.Simulator(
        seed=1234, tile_bits=args.tile_bits, coalesce_bits=args.coalesce_bits, mfma_blocks=not args.no_blocks,
        autotune=tuner or False, specialize_passes=not args.no_specialize,
    )
    t0 = time.perf_counter()
    comp = sim._compile(circuit, qubits)
    t_compile = time.perf_counter() - t0
    t_spec = comp.specialize_s
    _progress(f"{n}-qubit grid program compiled and tuned in {t_compile:.1f} s")
    tune_info = None
    if tuner is not None:
        import dataclasses

        from cirq_amd.autotune import default_config

        tune_info = {"config": dataclasses.asdict(comp.tuning), "compile_s": t_compile}
        if tuner.log:  # timed here (not read from the cache file)
            (log,) = tuner.log.values()
            times = dict(log)
            tune_info.update(candidates=len(log), default_ms=times[default_config(n, False)], tuned_ms=times[comp.tuning])
        else:
            tune_info["from_cache"] = os.environ.get("CIRQ_AMD_TUNE_CACHE")
        torch.cuda.empty_cache()  # the tuner's scratch state
    factors = product_state_factors(circuit, qubits)
    assert len(factors) == 1 and len(factors[0]) == n
    axes = [qmap[q] for q in factors[0]]
    dev = sim._dev()
    st = engine.new_state(n, dev)
    stream = torch.cuda.current_stream()
    pass_events = []

    def step(record: bool):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        comp.dev.run(st, basis=0)  # |0..0>: the first pass builds it instead of reading it
        e1.record(stream)
        if record:
            pass_events.append((e0, e1))
        u = sim._prng.random_sample(reps)
        idx = sim._sampler.draw(st, n, False, u, axes=axes)
        ev = engine.expect_terms(st, n, False, terms, sim._ws)  # queued behind the sampling kernels
        bits = sim._sampler.bits(idx, n)
        vals = ev.cpu().numpy()
        return bits, vals

    for _ in range(warmup):
        step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        bits, vals = step(True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    pass_ms = sum(a.elapsed_time(b) for a, b in pass_events) / len(pass_events)
    n_pass = int(comp.packed.n_passes)
    avg_launch_s = pass_ms / 1e3 / n_pass
    # every pass writes the state once and reads what the state can reach so far (the first builds
    # |0..0> and reads nothing): the average algorithmic bytes per launch
    bytes_per_launch = _chain_bytes(comp.packed, n) / n_pass
    achieved = bytes_per_launch / avg_launch_s / 1e9
    ms_per_step = wall / steps * 1e3
    layers_per_s = depth * steps / wall
    norm = engine.norm2(st, n, sim._ws)  # FP64 on the GPU: a size-independent check of the passes
    return dict(n=n, n_ops=n_ops, zz=zz, comp=comp, bits=bits, vals=vals, wall=wall, pass_ms=pass_ms, n_pass=n_pass,
                avg_launch_s=avg_launch_s, bytes_per_launch=bytes_per_launch, achieved=achieved,
                ms_per_step=ms_per_step, layers_per_s=layers_per_s, tune_info=tune_info, norm=norm,
                specialize_s=t_spec)



def q32_leg(args):
    """The larger configuration's 32-qubit variant: the random grid circuit on a 4x8 grid, depth 20
    (32 GiB complex64 state), with the same step as the main leg (10^6 samples, 52 <ZZ>)."""
    g = grid_workload(args, 4, 8, 20, args.reps, args.q32_steps, 1, autotune=not args.no_autotune)
    return {
        "workload": f"random grid circuit 4x8 (32 qubits), depth 20, {g['n_ops']} gates, {args.reps} samples, "
        f"{len(g['zz'])} <ZZ>",
        "layers_per_s": g["layers_per_s"],
        "ms_per_step": g["ms_per_step"],
        "steps": args.q32_steps,
        "passes": g["n_pass"],
        "passes_ms_per_step": g["pass_ms"],
        "pass_kernel_avg_ms": g["avg_launch_s"] * 1e3,
        "pass_hbm_gbs": g["achieved"],
        "pass_roofline_frac": g["achieved"] / HBM_PEAK_GBS,
        "norm_error": abs(g["norm"] - 1.0),
        "autotune": g["tune_info"],
        "specialized_passes": g["specialize_s"] is not None,
        "specialize_s": g["specialize_s"],
    }


LEGS = {"q32": lambda a: q32_leg(a), "adjoint": lambda a: adjoint_leg(a), "qaoa": lambda a: qaoa_leg(a),
        "dm_adjoint": lambda a: dm_adjoint_leg(a), "sweep": lambda a: sweep_leg(a),
        "trajectories": lambda a: trajectory_leg(a), "density_matrix": lambda a: dm_leg(a)}


def main():
    args = parse()
    if args.gpus != 1:
        if int(os.environ.get("RANK", "0")) == 0:
            print(
                "bench.py runs on one GPU only: this simulator has no multi-GPU path "
                "(a state would shard by its top qubit bits, one process per GPU)", file=sys.stderr
            )
        sys.exit(2)

    import torch

    if args.copy_calibration:
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": 4 * n, "copy_calibration_bytes_written": 4 * n, "copies": 5}))
        return

    if args.leg:
        print(json.dumps({args.leg: LEGS[args.leg](args)}), flush=True)
        return

    t0 = time.perf_counter()
    g = grid_workload(args, args.rows, args.cols, args.depth, args.reps, args.steps, args.warmup,
                      autotune=not args.no_autotune)
    _progress(f"main step done in {time.perf_counter() - t0:.1f} s")
    n, n_ops, zz, comp, bits = g["n"], g["n_ops"], g["zz"], g["comp"], g["bits"]
    wall, pass_ms, n_pass = g["wall"], g["pass_ms"], g["n_pass"]
    avg_launch_s, bytes_per_launch, achieved = g["avg_launch_s"], g["bytes_per_launch"], g["achieved"]
    ms_per_step, layers_per_s, tune_info = g["ms_per_step"], g["layers_per_s"], g["tune_info"]

    traffic = None  # measured HBM bytes per launch (rocprofv3 PMC), if measured for this tile width
    spec = g["specialize_s"] is not None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{'k_spec' if spec else 'k_pass'}_{n}q.json")
    tile_l = int(comp.packed.pass_arr["L"].max())
    if os.path.exists(tpath):
        with open(tpath) as f:
            rec = json.load(f)
        if rec.get("tile_bits") == tile_l and bool(rec.get("specialized")) == spec:
            traffic = rec.get("bytes_per_launch")

    out = {
        "metric": "30q layers/s & HBM GB/s; adjoint vs param-shift grads/s at grad err<=1e-5",
        "value": layers_per_s,
        "unit": "layers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex64 (fp32 pairs); fused 4-qubit blocks on MFMA with split-FP16 operands (hi + lo, three "
        "products accumulated in fp32, ~2^-22 relative per block); fp64 reductions",
        "data": "synthetic: seeded random grid circuit (Cirq's generator restated)",
        "config": {
            "workload": f"random grid circuit {args.rows}x{args.cols} ({n} qubits), depth {args.depth}, "
            f"{n_ops} gates, {args.reps} samples, {len(zz)} <ZZ>",
            "qubits": n,
            "depth": args.depth,
            "gates": n_ops,
            "samples": args.reps,
            "zz_pairs": len(zz),
            "passes": n_pass,
            "tile_bits": int(comp.packed.pass_arr["L"].max()) if n_pass else 0,
            "mfma_blocks": int((comp.packed.ops_arr["kind"] == 5).sum()),
            "specialized_passes": g["specialize_s"] is not None,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
        },
        "extra": {
            "pass_kernel_avg_ms": avg_launch_s * 1e3,
            "passes_ms_per_step": pass_ms,
            "gate_applications_per_s": n_ops * args.steps / wall,
            "hbm_gbs_whole_step": n_pass * bytes_per_launch / (ms_per_step / 1e3) / 1e9,
            "sample_check": int(bits.shape[0]),
            "norm_error": abs(g["norm"] - 1.0),
            "autotune": tune_info,
            "specialize_s": g["specialize_s"],
        },
    }
    legs = [("q32", args.no_q32), ("adjoint", args.no_adjoint), ("qaoa", args.no_qaoa),
            ("dm_adjoint", args.no_dm_adjoint), ("sweep", args.no_sweep), ("trajectories", args.no_trajectories),
            ("density_matrix", args.no_dm)]
    for name, skip in legs:
        if not skip:
            t0 = time.perf_counter()
            out["extra"][name] = LEGS[name](args)
            _progress(f"{name} leg done in {time.perf_counter() - t0:.1f} s")
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        out["cpu_baseline"] = cpu_baseline(args, args.rows, args.cols, args.depth, n)
        _progress(f"cpu baseline done in {time.perf_counter() - t0:.1f} s")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
