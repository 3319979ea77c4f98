"""Writes the oracle fixtures for the large-size GPU parity tests (run once on a CPU host).

* ``grid28_depth40.npz``: the random grid circuit on the first 28 qubits of a 5x6 grid, depth 40,
  seed 0, from |0...0> in complex128 (oracle/csim.c, pinned against the NumPy restatement by
  tests/test_oracle_csim.py).  Stored: 4,098 amplitudes at seeded indices (plus the first and last),
  <Z_i Z_j> for every nearest-neighbour pair, and the overlaps <v_r|psi> with 4 seeded product
  states |v_r> = (x)_q V_rq |0> (each V_rq a Haar-random 2x2 unitary, written down too).
* ``qaoa24_p8.npz``: QAOA MaxCut on the seeded 24-node 3-regular graph, p = 8 (16 symbols, seed 0):
  <H> and the gradient by central differences (h = 1e-5) of the complex128 simulation, accurate to
  about 1e-9.
* ``noisy_line12_depth20.npz``: the noisy-line density-matrix circuit on 12 qubits, 20 layers,
  seed 0 (depolarize(1e-3) after every gate, amplitude_damp(1e-3) per qubit per layer), from
  |0...0><0...0|.  Stored: the diagonal of rho, 4,096 off-diagonal entries at seeded (row, col)
  positions and Tr(rho sum_i Z_i).

Usage: ``python tests/golden/make_large_fixtures.py [grid|dm|qaoa|all]`` from the repository root.
The 28-qubit run holds 8 GiB and takes about 10 minutes on 8 cores.
"""

from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from cirq_amd import generators  # noqa: E402
from oracle import cirq_np as O  # noqa: E402
from oracle import csim  # noqa: E402
from tests.helpers import to_oracle_ops  # noqa: E402

GRID_N, GRID_DEPTH, GRID_SEED = 28, 40, 0
DM_N, DM_LAYERS, DM_SEED = 12, 20, 0


def haar_2x2(rng) -> np.ndarray:
    z = (rng.randn(2, 2) + 1j * rng.randn(2, 2)) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diag(r)
    return q * (d / np.abs(d))


def grid_fixture_inputs():
    """The seeded pieces both the fixture and the GPU test derive: circuit, qubits, indices, V."""
    c = generators.grid_circuit(5, 6, GRID_DEPTH, seed=GRID_SEED, n_qubits=GRID_N)
    qs = sorted(c.all_qubits())
    rng = np.random.RandomState(2800)
    idx = np.concatenate([[0, 2**GRID_N - 1], rng.randint(0, 2**GRID_N, 4096)]).astype(np.int64)
    V = np.array([[haar_2x2(rng) for _ in range(GRID_N)] for _ in range(4)])  # (4, n, 2, 2)
    return c, qs, idx, V


def make_grid(path: str):
    c, qs, idx, V = grid_fixture_inputs()
    ops = to_oracle_ops(c, qs)
    t0 = time.time()
    psi = csim.simulate_state_vector(ops, GRID_N)
    print(f"grid {GRID_N}q depth {GRID_DEPTH}: {len(ops)} ops in {time.time() - t0:.0f} s", flush=True)
    qmap = {q: i for i, q in enumerate(qs)}
    pairs = [tuple(qmap[q] for q in z.qubits) for z in generators.nearest_neighbour_zz(qs)]
    zmasks = [(1 << (GRID_N - 1 - a)) | (1 << (GRID_N - 1 - b)) for a, b in pairs]
    zz = csim.expect_z(psi, GRID_N, zmasks)
    amp = psi[idx]
    norm = float(np.vdot(psi, psi).real)
    overlaps = []
    for r in range(V.shape[0]):
        w = psi.copy()
        for a in range(GRID_N):
            csim.apply_matrix(w, GRID_N, V[r, a].conj().T, [a])
        overlaps.append(w[0])
        del w
    np.savez_compressed(path, idx=idx, amp=amp, pairs=np.array(pairs, dtype=np.int32), zz=zz, V=V,
                        overlaps=np.array(overlaps), norm=norm, n=GRID_N, depth=GRID_DEPTH, seed=GRID_SEED)
    print(f"wrote {path}: norm {norm:.15f}, |overlaps| {np.abs(overlaps)}", flush=True)


def dm_fixture_inputs():
    nc = generators.noisy_line_circuit(DM_N, DM_LAYERS, seed=DM_SEED)
    qs = sorted(nc.all_qubits())
    rng = np.random.RandomState(1200)
    rows = rng.randint(0, 2**DM_N, 4096).astype(np.int64)
    cols = rng.randint(0, 2**DM_N, 4096).astype(np.int64)
    return nc, qs, rows, cols


def make_dm(path: str):
    nc, qs, rows, cols = dm_fixture_inputs()
    ops = to_oracle_ops(nc, qs)
    t0 = time.time()
    rho = csim.simulate_density_matrix(ops, DM_N).reshape(2**DM_N, 2**DM_N)
    print(f"noisy line {DM_N}q x {DM_LAYERS}: {len(ops)} ops in {time.time() - t0:.0f} s", flush=True)
    diag = np.real(np.diag(rho)).copy()
    ent = rho[rows, cols].copy()
    sz = float(sum(O.expectation_from_density_matrix(rho, (1.0, {i: "Z"}), DM_N) for i in range(DM_N))) \
        if DM_N <= 8 else float(sum(np.sum(diag * (1 - 2 * ((np.arange(2**DM_N) >> (DM_N - 1 - i)) & 1)))
                                    for i in range(DM_N)))
    np.savez_compressed(path, diag=diag, rows=rows, cols=cols, entries=ent, tr_rho_sum_z=sz,
                        trace=float(np.real(np.trace(rho))), n=DM_N, layers=DM_LAYERS, seed=DM_SEED)
    print(f"wrote {path}: Tr(rho sum Z) = {sz:.12f}, trace {np.real(np.trace(rho)):.15f}", flush=True)


def make_qaoa(path: str, n: int = 24, p: int = 8, seed: int = 0, h: float = 1e-5):
    from cirq_amd.study import ParamResolver

    circuit, syms, vals, ham, edges = generators.qaoa_maxcut(n, p, seed=seed)
    qs = sorted(circuit.all_qubits())
    qmap = {q: i for i, q in enumerate(qs)}
    const = 0.0
    masks, coefs = [], []
    for t in ham:
        items = list(t.items())
        if not items:
            const += complex(t.coefficient).real
            continue
        masks.append(sum(1 << (n - 1 - qmap[q]) for q, _ in items))
        coefs.append(complex(t.coefficient).real)

    def value(v):
        psi = csim.simulate_state_vector(to_oracle_ops(circuit, qs, ParamResolver(dict(zip(syms, v)))), n)
        return const + float(np.dot(coefs, csim.expect_z(psi, n, masks)))

    t0 = time.time()
    v0 = value(vals)
    grad = np.zeros(len(syms))
    for k in range(len(syms)):
        up, dn = vals.copy(), vals.copy()
        up[k] += h
        dn[k] -= h
        grad[k] = (value(up) - value(dn)) / (2 * h)
        print(f"  symbol {k}: {grad[k]:+.10f} ({time.time() - t0:.0f} s)", flush=True)
    np.savez_compressed(path, value=v0, grad=grad, values=vals, symbols=np.array([s.name for s in syms]),
                        edges=np.array(edges), n=n, p=p, seed=seed, h=h)
    print(f"wrote {path}: <H> = {v0:.12f}", flush=True)


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    if what in ("dm", "all"):
        make_dm(os.path.join(HERE, "noisy_line12_depth20.npz"))
    if what in ("grid", "all"):
        make_grid(os.path.join(HERE, "grid28_depth40.npz"))
    if what in ("qaoa", "all"):
        make_qaoa(os.path.join(HERE, "qaoa24_p8.npz"))
