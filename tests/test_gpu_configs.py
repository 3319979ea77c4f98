"""GPU parity at the benchmark configurations' own sizes.

* Random grid circuit, 28 qubits (first 28 of 5x6), depth 40, compiled as the 30-qubit benchmark
  compiles it (13-bit tiles, 4 coalescing bits, blocks of up to 4 qubits on MFMA, fused ops that
  depend on up to 3 outer bits) against the complex128 oracle fixture
  (tests/golden/grid28_depth40.npz): 4,098 amplitudes, 48 <Z_i Z_j> and 4 product-state overlaps.
  Bars: sampled fidelity >= 1 - 1e-6, max-abs amplitude error <= 1e-5, <ZZ> within 1e-5.
* The same compiled program's U then U^-1 at 30 qubits (depth 40) and 32 qubits (4x8, depth 20):
  |<0|U^-1 U|0>|^2 / <psi|psi> >= 1 - 1e-6, and the FP64 norm after U within 1e-5 of 1.
* Sampling bit-exact with NumPy's choice at 26 and 28 qubits (2^26 / 2^28 bins, the multi-
  thousand-superchunk walk), and at draws placed exactly on NumPy's cdf values and their
  neighbouring doubles for adversarial probability vectors (1 ulp of the float64 cumsum decides).
* Noisy density matrix, 12-qubit line x 20 layers, against the oracle fixture
  (tests/golden/noisy_line12_depth20.npz); the 16-qubit line x 20 layers (2^32 entries) by
  properties (trace 1, Hermitian, real non-negative diagonal, purity <= 1, Tr(rho sum Z) equal
  from two kernels).
* QAOA MaxCut, 24-node 3-regular graph, p = 8 (16 symbols): adjoint vs GPU parameter shift; oracle
  finite differences at 10 nodes, p = 8.
* Mixed-size sweep, 4,096 points over 256 circuits of 10-22 qubits: the batched path against the
  single-state path on one point of every circuit.
"""

import os
import sys

import numpy as np
import pytest
import torch

import cirq_amd as cirq
from cirq_amd import engine, generators, tape
from cirq_amd.autotune import TuneConfig
from cirq_amd.paulis import pauli_masks

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# the settings the autotuner picks for the 30- and 32-qubit grid circuits
BENCH_TUNING = TuneConfig(tile_bits=13, coalesce_bits=4, kmax=4, variant_bits=3, mfma_weight=1.0)


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _program(circuit, qubits, cfg=BENCH_TUNING, dev=None):
    low = tape.Lowering(tuple(qubits), False)
    low.lower_circuit(circuit)
    n = len(qubits)
    packed = tape.compile_ops(low.ops, n, **cfg.compile_kwargs(n))
    return packed, engine.DeviceProgram(packed, dev or engine.require_gpu())


def _sample_fidelity(got, want):
    return abs(np.vdot(want, got)) ** 2 / (np.vdot(want, want).real * np.vdot(got, got).real)


# ------------------------------------------------------------------------------ 28-qubit grid vs the oracle
@pytest.fixture(scope="module")
def grid28(gpu):
    from tests.golden.make_large_fixtures import grid_fixture_inputs

    fx = np.load(os.path.join(GOLDEN, "grid28_depth40.npz"))
    c, qs, idx, V = grid_fixture_inputs()
    assert np.array_equal(idx, fx["idx"]) and np.allclose(V, fx["V"])
    _log(f"grid28: compiling {sum(1 for _ in c.all_operations())} gates with the benchmark tuning")
    packed, prog = _program(c, qs)
    assert int(packed.pass_arr["L"].max()) == 13 and int(((packed.ops_arr["kind"] & 0xFF) == 5).sum()) > 0
    st = engine.empty_state(28, gpu)
    prog.run(st, basis=0)
    torch.cuda.synchronize()
    _log(f"grid28: {int(packed.n_passes)} passes done")
    yield dict(c=c, qs=qs, idx=idx, V=V, fx=fx, st=st)
    del st
    torch.cuda.empty_cache()


def test_grid28_tuned_program_matches_oracle_fixture(grid28):
    st, fx, qs, V = grid28["st"], grid28["fx"], grid28["qs"], grid28["V"]
    got = st[torch.from_numpy(grid28["idx"]).to(st.device)].cpu().numpy().astype(np.complex128)
    want = fx["amp"]
    err = np.abs(got - want)
    fid = _sample_fidelity(got, want)
    rel = np.linalg.norm(got - want) / np.linalg.norm(want)
    _log(f"grid28 amplitudes: max|err| {err.max():.3e} (max|a| {np.abs(want).max():.3e}), rel L2 {rel:.3e}, "
         f"sampled fidelity 1-{1 - fid:.3e}")
    assert err.max() <= 1e-5
    assert fid >= 1 - 1e-6
    # <Z_i Z_j> over the nearest-neighbour pairs (FP64 reductions on the GPU)
    n = 28
    terms = [(0, (1 << (n - 1 - int(a))) | (1 << (n - 1 - int(b))), 1.0) for a, b in fx["pairs"]]
    zz = engine.expect_terms(st, n, False, terms, engine.Workspace(st.device)).cpu().numpy().real
    _log(f"grid28 <ZZ>: max|err| {np.abs(zz - fx['zz']).max():.3e}")
    assert np.abs(zz - fx["zz"]).max() < 1e-5
    assert abs(engine.norm2(st, n, engine.Workspace(st.device)) - 1) < 1e-5
    # <v_r|psi> for product states: U_r^dagger layer on a copy, amplitude 0
    for r in range(V.shape[0]):
        layer = cirq.Circuit([cirq.MatrixGate(V[r, a].conj().T).on(q) for a, q in enumerate(qs)])
        _, lp = _program(layer, qs, TuneConfig(tile_bits=12, coalesce_bits=4, kmax=4))
        w = st.clone()
        lp.run(w)
        ov = complex(w[0].item())
        del w
        d = abs(ov - fx["overlaps"][r])
        _log(f"grid28 overlap {r}: |ref| {abs(fx['overlaps'][r]):.3e}, |diff| {d:.3e}")
        assert d <= 1e-4 * abs(fx["overlaps"][r]) + 1e-9


def test_grid28_sampling_bit_exact(grid28):
    """2^28 bins: NumPy's choice on the float32 probabilities of the GPU state vs the GPU sampler."""
    from oracle import cirq_np as O

    st = grid28["st"]
    sim = cirq.Simulator()
    sim._dev()
    reps = 100_000
    u = np.random.RandomState(28).random_sample(reps)
    got = sim._sampler.draw(st, 28, False, u).cpu().numpy()
    psi = st.cpu().numpy()
    p = O.state_vector_probs(psi, list(range(28)))
    del psi
    want = np.random.RandomState(28).choice(len(p), size=reps, p=p)
    _log(f"grid28 sampling: {int(np.sum(got != want))} of {reps} differ")
    assert np.array_equal(got, want)


def test_grid26_sampling_bit_exact(gpu):
    from oracle import cirq_np as O

    c = generators.grid_circuit(5, 6, 16, seed=3, n_qubits=26)
    sim = cirq.Simulator(seed=0)
    st = sim.simulate(c)._state
    reps = 200_000
    u = np.random.RandomState(26).random_sample(reps)
    got = sim._sampler.draw(st, 26, False, u).cpu().numpy()
    p = O.state_vector_probs(st.cpu().numpy(), list(range(26)))
    want = np.random.RandomState(26).choice(len(p), size=reps, p=p)
    _log(f"grid26 sampling: {int(np.sum(got != want))} of {reps} differ")
    assert np.array_equal(got, want)


def _numpy_cdf(q32: np.ndarray) -> np.ndarray:
    """RandomState.choice's cdf for p = q / q.sum() in float32 (Cirq's normalisation)."""
    p = q32 / q32.sum()
    cdf = p.astype(np.float64).cumsum()
    cdf /= cdf[-1]
    return cdf


@pytest.mark.parametrize("n,kind", [(22, "binades"), (23, "spikes"), (24, "binades")])
def test_sampling_at_cdf_boundaries(gpu, n, kind):
    """Draws exactly at NumPy's cdf values and the doubles on either side: searchsorted(side='right')
    flips at these points, so any 1-ulp difference in the GPU's float64 cumsum changes an index."""
    rng = np.random.RandomState(n)
    N = 1 << n
    if kind == "binades":  # magnitudes spread over 40 binades, so the running sum crosses binades inside chunks
        q = (rng.exponential(size=N) * np.exp2(rng.randint(-40, 0, size=N))).astype(np.float32)
        q[rng.randint(0, N, N // 64)] = 0
    else:  # mostly tiny values with a few large spikes
        q = (rng.uniform(size=N) * 1e-9).astype(np.float32)
        q[rng.randint(0, N, 64)] = rng.uniform(0.01, 1, 64).astype(np.float32)
    cdf = _numpy_cdf(q)
    picks = np.unique(np.concatenate([
        rng.randint(0, N - 1, 3000),
        np.arange(255, N - 1, 1 << 14), np.arange(8191, N - 1, 1 << 16), np.arange((1 << 18) - 1, N - 1, 1 << 18),
    ]))
    base = cdf[picks]
    u = np.concatenate([base, np.nextafter(base, 0.0), np.nextafter(base, 1.0)])
    u = u[(u >= 0) & (u < 1)]
    want = np.searchsorted(cdf, u, side="right")
    sim = cirq.Simulator()
    sim._dev()
    probs = torch.from_numpy(q).to(gpu)
    got = sim._sampler.draw(None, n, False, u, probs=probs).cpu().numpy()
    _log(f"cdf boundaries n={n} {kind}: {len(u)} draws, {int(np.sum(got != want))} differ")
    assert np.array_equal(got, want)


# ------------------------------------------------------------------------------ 30- and 32-qubit round trips
@pytest.mark.parametrize("rows,cols,depth", [(5, 6, 40), (4, 8, 20)])
def test_grid_round_trip_fidelity(gpu, rows, cols, depth):
    n = rows * cols
    _log(f"round trip {n}q depth {depth}: compiling U and U^-1")
    c = generators.grid_circuit(rows, cols, depth, seed=0)
    qs = sorted(c.all_qubits())
    inv = cirq.Circuit([cirq.MatrixGate(np.conj(op.gate._unitary_()).T).on(*op.qubits)
                        for op in reversed(list(c.all_operations()))])
    pu, u = _program(c, qs)
    pv, v = _program(inv, qs)
    assert int(pu.pass_arr["L"].max()) == 13
    st = engine.empty_state(n, gpu)
    ws = engine.Workspace(gpu)
    u.run(st, basis=0)
    n2 = engine.norm2(st, n, ws)
    v.run(st)
    a0 = complex(st[0].item())
    n2b = engine.norm2(st, n, ws)
    fid = abs(a0) ** 2 / n2b
    _log(f"round trip {n}q: norm^2 after U 1{n2 - 1:+.3e}, after U^-1 U 1{n2b - 1:+.3e}, |a0|^2 1{abs(a0) ** 2 - 1:+.3e}, "
         f"fidelity 1-{1 - fid:.3e}")
    del st
    torch.cuda.empty_cache()
    assert abs(n2 - 1) < 1e-5
    assert fid >= 1 - 1e-6


# ------------------------------------------------------------------------------ noisy density matrix
def test_noisy_line12_matches_oracle_fixture(gpu):
    from tests.golden.make_large_fixtures import dm_fixture_inputs

    fx = np.load(os.path.join(GOLDEN, "noisy_line12_depth20.npz"))
    nc, qs, rows, cols = dm_fixture_inputs()
    assert np.array_equal(rows, fx["rows"]) and np.array_equal(cols, fx["cols"])
    sim = cirq.DensityMatrixSimulator()
    comp = sim._compile(nc, qs)
    rho = engine.empty_state(24, sim._dev())
    comp.dev.run(rho, basis=0)
    m = rho.view(4096, 4096)
    diag = m.diagonal().cpu().numpy()
    ent = m[torch.from_numpy(rows).to(rho.device), torch.from_numpy(cols).to(rho.device)].cpu().numpy()
    ev = sim.simulate_expectation_values(nc, [generators.sum_z(qs)])[0]
    d_err = np.abs(diag - fx["diag"]).max()
    e_err = np.abs(ent - fx["entries"]).max()
    _log(f"noisy line 12q: diag max|err| {d_err:.3e} (sum {np.abs(diag - fx['diag']).sum():.3e}), entries {e_err:.3e}, "
         f"Tr(rho sum Z) {ev:.9f} vs {float(fx['tr_rho_sum_z']):.9f}")
    assert np.abs(diag.imag).max() < 1e-6
    # Frobenius bar 1e-5 over the whole matrix; on the 4,096 diagonal entries and 4,096 samples
    assert np.linalg.norm(diag - fx["diag"]) < 1e-5
    assert e_err < 1e-6
    assert abs(ev - float(fx["tr_rho_sum_z"])) < 1e-5


def test_noisy_line16_properties(gpu):
    """2^32 entries (32 GiB): trace, Hermiticity, diagonal and purity checked on the GPU."""
    n = 16
    nc = generators.noisy_line_circuit(n, 20, seed=0)
    qs = sorted(nc.all_qubits())
    sim = cirq.DensityMatrixSimulator()
    comp = sim._compile(nc, qs)
    rho = engine.empty_state(2 * n, sim._dev())
    _log("noisy line 16q: passes")
    comp.dev.run(rho, basis=0)
    qmap = {q: i for i, q in enumerate(qs)}
    terms = [pauli_masks(t, qmap, n) for t in generators.sum_z(qs)]
    ev = engine.expect_terms(rho, n, True, terms, sim._ws).sum().real.item()
    m = rho.view(1 << n, 1 << n)
    d = m.diagonal()
    tr = d.real.double().sum().item()
    idx = torch.arange(1 << n, device=rho.device)
    signs = torch.zeros(1 << n, dtype=torch.float64, device=rho.device)
    for i in range(n):
        signs += 1 - 2 * ((idx >> (n - 1 - i)) & 1).double()
    ev_diag = (d.real.double() * signs).sum().item()
    rng = np.random.RandomState(16)
    r = torch.from_numpy(rng.randint(0, 1 << n, 20000)).to(rho.device)
    c = torch.from_numpy(rng.randint(0, 1 << n, 20000)).to(rho.device)
    herm = (m[r, c] - m[c, r].conj()).abs().max().item()
    pur = 0.0
    for k in range(0, 1 << n, 4096):  # Tr(rho^2) = sum |rho_ij|^2 for Hermitian rho
        pur += m[k:k + 4096].abs().double().pow(2).sum().item()
    _log(f"noisy line 16q: trace 1{tr - 1:+.3e}, Tr(rho sum Z) {ev:.9f} (diag {ev_diag:.9f}), "
         f"max|rho_ij - conj rho_ji| {herm:.3e}, min diag {d.real.min().item():.3e}, max|Im diag| "
         f"{d.imag.abs().max().item():.3e}, purity {pur:.6f}")
    del rho, m, d
    torch.cuda.empty_cache()
    assert abs(tr - 1) < 1e-5
    assert abs(ev - ev_diag) < 1e-6
    assert herm < 1e-6
    assert pur <= 1 + 1e-5


# ------------------------------------------------------------------------------ QAOA 24 nodes, p = 8
def test_qaoa24_p8_adjoint_vs_parameter_shift(gpu):
    from cirq_amd.adjoint import AdjointGradient

    circuit, syms, vals, ham, edges = generators.qaoa_maxcut(24, 8, seed=0)
    assert len(edges) == 36 and len(syms) == 16
    ag = AdjointGradient(circuit, ham)
    value, grad = ag.value_and_grad(dict(zip(syms, vals)))
    v2, g2 = ag.parameter_shift(dict(zip(syms, vals)))
    _log(f"QAOA 24q p=8: <H> {value:.6f} vs {v2:.6f}, max|adjoint - shift| {np.abs(grad - g2).max():.3e}")
    # complex64 amplitudes: errors scale with the observable's norm (sum |c_t| = 36)
    assert abs(value - v2) < 1e-5 * 36
    np.testing.assert_allclose(grad, g2, atol=1e-5 * 36 / 10)


def test_qaoa24_p8_adjoint_matches_oracle_fixture(gpu):
    """The 24-node p = 8 gradient against central differences of the complex128 simulation
    (tests/golden/qaoa24_p8.npz, accurate to ~1e-8).  Bar: 1e-5 relative to the largest gradient
    component (12.9 here), i.e. 1.3e-4 absolute.  complex64 states after ~1,000 gate applications
    (forward sweep and un-computation) carry ~2e-6 relative error, and these gradients sum 36-60
    such terms of size up to |H| = 18 (the identity part of H is dropped from phi); measured:
    2.5e-5 absolute."""
    from cirq_amd.adjoint import AdjointGradient

    fx = np.load(os.path.join(GOLDEN, "qaoa24_p8.npz"))
    circuit, syms, vals, ham, edges = generators.qaoa_maxcut(24, 8, seed=0)
    assert np.array_equal(vals, fx["values"]) and [s.name for s in syms] == list(fx["symbols"])
    ag = AdjointGradient(circuit, ham)
    value, grad = ag.value_and_grad(dict(zip(syms, vals)))
    want = dict(zip(syms, fx["grad"]))
    err = np.abs(np.asarray(grad) - np.array([want[s] for s in ag.symbols]))
    _log(f"QAOA 24q p=8 vs oracle: max|grad err| {err.max():.3e} (max|grad| {np.abs(fx['grad']).max():.2f}), "
         f"<H> err {abs(value - float(fx['value'])):.3e}")
    assert err.max() <= 1e-5 * max(1.0, np.abs(fx["grad"]).max())
    assert abs(value - float(fx["value"])) < 1e-5


def test_qaoa10_p8_matches_oracle_finite_differences(gpu):
    from cirq_amd.adjoint import AdjointGradient
    from tests.test_adjoint_host import _oracle_ev, _oracle_grad

    circuit, syms, vals, ham, _ = generators.qaoa_maxcut(10, 8, seed=2)
    qubits = sorted(circuit.all_qubits())
    ag = AdjointGradient(circuit, ham)
    value, grad = ag.value_and_grad(dict(zip(syms, vals)))
    want = dict(zip(syms, _oracle_grad(circuit, qubits, ham, syms, vals)))
    err = np.abs(np.asarray(grad) - np.array([want[s] for s in ag.symbols])).max()
    _log(f"QAOA 10q p=8 vs oracle finite differences: max err {err:.3e}")
    assert err <= 1e-5 * 15 / 10
    assert abs(value - _oracle_ev(circuit, qubits, ham, cirq.ParamResolver(dict(zip(syms, vals))))) < 1e-5 * 15


# ------------------------------------------------------------------------------ 4,096-point mixed sweep
def test_full_mixed_sweep_against_single_state_path(gpu):
    from cirq_amd.adjoint import AdjointGradient
    from cirq_amd.batch import BatchedExpectation
    from cirq_amd.paulis import PauliSum

    circuits, observables, values = generators.mixed_sweep_workload(n_circuits=256, points=16, seed=0)
    assert sum(v.shape[0] for v in values) == 4096
    _log("mixed sweep: compiling 256 circuits")
    be = BatchedExpectation(circuits, observables)
    evs, grads = be.run(values)
    _log("mixed sweep: batched run done; single-state path on one point per circuit")
    worst = 0.0
    for c in range(len(circuits)):
        b = c % values[c].shape[0]
        ag = AdjointGradient(circuits[c], PauliSum.from_pauli_strings(observables[c]))
        v, g = ag.value_and_grad(values[c][b])
        worst = max(worst, float(np.abs(g - grads[c][b]).max()), abs(v - float(evs[c][b].sum())))
        del ag
    _log(f"mixed sweep: max |batched - single| {worst:.3e}")
    assert worst < 1e-5
