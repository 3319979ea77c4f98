"""GPU parity tests: the HIP engine against the NumPy oracle and NumPy's own sampling.

Tolerances (complex64 engine vs complex128 oracle): max-abs amplitude error <= 1e-5 and fidelity
>= 1 - 1e-6; density matrices within 1e-5 Frobenius; samples bit-exact given the same uniforms.
"""

import sys

import numpy as np
import pytest
import torch

import cirq_amd as cirq
from cirq_amd import engine, generators
from oracle import cirq_np as O
from tests.helpers import fidelity, terms_of, to_oracle_ops
from tests.test_host import _random_circuit

pytestmark = pytest.mark.gpu

ATOL = 1e-5


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("tile", [(3, 1), (5, 2), (8, 4), (10, 4)])
def test_random_circuits_match_oracle(gpu, seed, tile):
    _log(f"random circuit seed={seed} tile={tile}")
    rng = np.random.RandomState(seed)
    n = 10
    c, qs = _random_circuit(n, 120, rng)
    sim = cirq.Simulator(tile_bits=tile[0], coalesce_bits=tile[1])
    got = sim.simulate(c, qubit_order=qs).final_state_vector
    want = O.simulate_state_vector(to_oracle_ops(c, qs), n)
    assert np.max(np.abs(got - want)) <= ATOL
    assert fidelity(got, want) >= 1 - 1e-6


@pytest.mark.parametrize("n,depth,seed", [(12, 12, 0), (16, 14, 1), (20, 20, 0)])
def test_grid_circuit_matches_oracle(gpu, n, depth, seed):
    _log(f"grid circuit n={n} depth={depth}")
    rows = {12: (3, 4), 16: (4, 4), 20: (4, 5)}[n]
    c = generators.grid_circuit(*rows, depth, seed=seed)
    qs = sorted(c.all_qubits())
    got = cirq.Simulator().simulate(c).final_state_vector
    want = O.simulate_state_vector(to_oracle_ops(c, qs), n)
    assert np.max(np.abs(got - want)) <= ATOL
    assert fidelity(got, want) >= 1 - 1e-6


@pytest.mark.parametrize("tile_bits", [12, 13, 14])
@pytest.mark.parametrize("kind", ["grid", "random"])
def test_wide_tiles_match_oracle(gpu, tile_bits, kind):
    """Tiles of 12 and 13 bits (16 elements per thread, 4 register bits) on a 15-qubit state: the
    widths the large-state simulators run, with PERMUTE re-layouts and MFMA blocks in each pass."""
    _log(f"wide tiles tile_bits={tile_bits} kind={kind}")
    if kind == "grid":
        c = generators.grid_circuit(3, 5, 16, seed=4)
        qs = sorted(c.all_qubits())
    else:
        c, qs = _random_circuit(15, 200, np.random.RandomState(21))
    sim = cirq.Simulator(tile_bits=tile_bits, coalesce_bits=4)
    got = sim.simulate(c, qubit_order=qs).final_state_vector
    want = O.simulate_state_vector(to_oracle_ops(c, qs), len(qs))
    assert np.max(np.abs(got - want)) <= ATOL
    assert fidelity(got, want) >= 1 - 1e-6


def test_sampling_bit_exact_with_numpy_choice(gpu):
    """Same p-vector (from the GPU state) + same RandomState -> identical indices."""
    _log("sampling bit-exact")
    for n, seed in [(1, 1234), (3, 5), (8, 7), (14, 11), (18, 3), (21, 2)]:
        c = generators.grid_circuit(3, 7, 8, seed=seed, n_qubits=n) if n > 3 else cirq.Circuit(
            cirq.H.on_each(*cirq.LineQubit.range(n)), cirq.X(cirq.LineQubit(0)) ** 0.3
        )
        sim = cirq.Simulator()
        res = sim.simulate(c)
        st = res._state
        psi = st.cpu().numpy()
        reps = 20000
        p = O.state_vector_probs(psi, list(range(n)))
        want = np.random.RandomState(seed).choice(len(p), size=reps, p=p)
        u = np.random.RandomState(seed).random_sample(reps)
        idx = sim._sampler.draw(st, n, False, u).cpu().numpy()
        assert np.array_equal(idx, want), f"n={n}: {np.sum(idx != want)} mismatches"
        bits = sim._sampler.bits(torch.from_numpy(idx).to(st.device), n)
        want_bits = np.array([O.big_endian_int_to_digits(int(i), (2,) * n) for i in want[:200]])
        assert np.array_equal(bits[:200], want_bits)


def test_permuted_probabilities_and_sampling(gpu):
    """cqs_permuted_probs equals NumPy's float32 p of the transposed state, bit for bit, and
    sampling every qubit in a permuted order matches numpy's choice on that p."""
    _log("permuted probabilities")
    rng = np.random.RandomState(9)
    for n, seed in [(5, 1), (13, 2), (17, 3), (20, 4)]:
        c = generators.grid_circuit(4, 5, 6, seed=seed, n_qubits=n)
        sim = cirq.Simulator()
        st = sim.simulate(c)._state
        psi = st.cpu().numpy()
        # n = 20: the reversed order, whose tiles hold the lowest and the highest bits
        axes = list(range(n))[::-1] if n == 20 else list(rng.permutation(n))
        want_p = (psi.reshape((2,) * n).transpose(axes).reshape(-1) * np.conj(
            psi.reshape((2,) * n).transpose(axes).reshape(-1))).real.astype(np.float32)
        got_p = sim._sampler.permuted_probs(st, n, False, axes).cpu().numpy()
        assert np.array_equal(got_p, want_p), f"n={n}: {np.sum(got_p != want_p)} probabilities differ"
        p = want_p.astype(np.float64)
        p = p / p.sum()
        reps = 5000
        want = np.random.RandomState(seed).choice(len(p), size=reps, p=want_p / want_p.sum())
        u = np.random.RandomState(seed).random_sample(reps)
        idx = sim._sampler.draw(st, n, False, u, axes=axes).cpu().numpy()
        assert np.array_equal(idx, want)
    # density-matrix diagonal
    c = cirq.Circuit(cirq.H.on_each(*cirq.LineQubit.range(4)), cirq.CZ(cirq.LineQubit(0), cirq.LineQubit(3)),
                     cirq.X(cirq.LineQubit(1)) ** 0.37, cirq.depolarize(0.1).on(cirq.LineQubit(2)))
    dsim = cirq.DensityMatrixSimulator()
    rho_t = dsim.simulate(c)._state
    rho = rho_t.cpu().numpy().reshape(16, 16)
    axes = [2, 0, 3, 1]
    diag = np.real(np.diag(rho)).astype(np.float32).reshape((2,) * 4).transpose(axes).reshape(-1)
    got = dsim._sampler.permuted_probs(rho_t, 4, True, axes).cpu().numpy()
    assert np.array_equal(got, diag)


def test_marginal_sampling_bit_exact(gpu):
    _log("marginal sampling")
    c = generators.grid_circuit(3, 4, 6, seed=4)
    sim = cirq.Simulator()
    st = sim.simulate(c)._state
    psi = st.cpu().numpy()
    for axes in ([3, 0, 7], [11, 2], [5], list(range(12))[::-1]):
        p = O.state_vector_probs(psi, axes)
        probs = sim._sampler.marginal(st, 12, False, axes)
        raw = (psi * psi.conj()).real
        t = np.transpose(raw.reshape((2,) * 12), axes + [a for a in range(12) if a not in axes]).reshape(2 ** len(axes), -1)
        assert np.array_equal(probs.cpu().numpy(), np.sum(t, axis=-1)), axes
        u = np.random.RandomState(9).random_sample(5000)
        want = np.random.RandomState(9).choice(len(p), size=5000, p=p)
        got = sim._sampler.draw(None, len(axes), False, u, probs=probs).cpu().numpy()
        assert np.array_equal(got, want), axes


def test_run_seeded_known_answer(gpu):
    # cirq-core/cirq/sim/sparse_simulator_test.py:1046-1110
    a = cirq.NamedQubit("a")
    circuit = cirq.Circuit(cirq.X(a) ** 0.5, cirq.measure(a, key="a"))
    sim = cirq.Simulator(seed=1234)
    r1 = sim.run(circuit, repetitions=30).measurements["a"][:, 0]
    r2 = sim.run(circuit, repetitions=30).measurements["a"][:, 0]
    assert "".join(map(str, r1)) == "010110011101110110110100110101"
    assert "".join(map(str, r2)) == "101011010100011101010110111111"


def test_run_matches_oracle_sampling_order(gpu):
    """Entangled circuit: one factor sampled in Cirq's merge order, same prng stream."""
    _log("run vs oracle")
    qs = cirq.LineQubit.range(5)
    c = cirq.Circuit(
        cirq.H(qs[4]), cirq.CNOT(qs[4], qs[1]), cirq.ry(0.7).on(qs[2]), cirq.CZ(qs[2], qs[1]),
        cirq.X(qs[0]) ** 0.4, cirq.CNOT(qs[0], qs[3]), cirq.ISWAP(qs[3], qs[2]) ** 0.5,
        cirq.measure(*qs, key="m"),
    )
    sim = cirq.Simulator(seed=77)
    got = sim.run(c, repetitions=2000).measurements["m"]
    # the oracle: factor order from Cirq's merge rule, probabilities from the GPU state
    from cirq_amd.simulator import product_state_factors, _split_terminal_measurements

    body, _ = _split_terminal_measurements(c)
    factors = product_state_factors(body, tuple(qs))
    assert len(factors) == 1
    order = factors[0]
    psi = cirq.Simulator().simulate(body, qubit_order=qs)._state.cpu().numpy()
    axes = [qs.index(q) for q in order]
    s = O.sample_state_vector(psi, axes, 2000, np.random.RandomState(77))
    want = s[:, [order.index(q) for q in qs]]
    assert np.array_equal(got, want)


def test_expectations_match_oracle(gpu):
    _log("expectations")
    c = generators.grid_circuit(3, 4, 8, seed=1)
    qs = sorted(c.all_qubits())
    psum = generators.random_pauli_sum(qs, 20, 3, seed=2)
    zz = generators.nearest_neighbour_zz(qs)
    sim = cirq.Simulator()
    got = sim.simulate_expectation_values(c, [psum] + zz)
    psi = O.simulate_state_vector(to_oracle_ops(c, qs), 12)
    want = [sum(O.expectation_from_state_vector(psi, t, 12) for t in terms_of(psum, qs))]
    want += [sum(O.expectation_from_state_vector(psi, t, 12) for t in terms_of(cirq.PauliSum.wrap(z), qs)) for z in zz]
    assert np.max(np.abs(np.array(got) - np.array(want))) < 1e-5


@pytest.mark.parametrize("n", [4, 5, 9, 16, 21])
def test_z_only_expectations_exact(gpu, n):
    """Terms without X/Y take the FP64 sign-split kernel: compare with an FP64 NumPy sum of the
    same complex64 state (40 random Z strings, more than one launch's worth, plus the norm)."""
    _log(f"Z-only expectations n={n}")
    rng = np.random.RandomState(n)
    psi = (rng.randn(2**n) + 1j * rng.randn(2**n)).astype(np.complex64)
    psi /= np.linalg.norm(psi)
    st = torch.from_numpy(psi).to(gpu)
    zs = [int(rng.randint(0, 2**n)) for _ in range(40)] + [0]
    coefs = rng.randn(len(zs)) + 1j * rng.randn(len(zs))
    terms = [(0, z, c) for z, c in zip(zs, coefs)]
    got = engine.expect_terms(st, n, False, terms, engine.Workspace(gpu)).cpu().numpy()
    p = np.abs(psi.astype(np.complex128)) ** 2
    idx = np.arange(2**n)
    for k, (z, c) in enumerate(zip(zs, coefs)):
        sign = 1 - 2 * (np.array([bin(v).count("1") for v in (idx & z)]) & 1)
        want = c * np.sum(sign * p)
        assert abs(got[k] - want) < 1e-12 * max(1.0, abs(c)) + 1e-13
    assert abs(engine.norm2(st, n, engine.Workspace(gpu)) - np.sum(p)) < 1e-12


def test_density_matrix_matches_oracle(gpu):
    _log("density matrix")
    nc = generators.noisy_line_circuit(5, 4, seed=3, p=0.02, gamma=0.03)
    qs = sorted(nc.all_qubits())
    sim = cirq.DensityMatrixSimulator()
    got = sim.simulate(nc).final_density_matrix
    want = O.simulate_density_matrix(to_oracle_ops(nc, qs), 5)
    assert np.linalg.norm(got - want) < 1e-5
    ev = sim.simulate_expectation_values(nc, [generators.sum_z(qs)])[0]
    want_ev = sum(O.expectation_from_density_matrix(want, t, 5) for t in terms_of(generators.sum_z(qs), qs))
    assert isinstance(ev, float) and abs(ev - want_ev) < 1e-5


def test_noisy_line_density_matrix_8_qubits(gpu):
    """The noisy-line density-matrix workload (Rz, Ry, alternating CZ, depolarize after every gate,
    amplitude damping each layer) at 8 qubits x 20 layers: a 16-bit index over several passes."""
    _log("noisy line 8q x 20")
    nc = generators.noisy_line_circuit(8, 20, seed=0)
    qs = sorted(nc.all_qubits())
    sim = cirq.DensityMatrixSimulator()
    got = sim.simulate(nc).final_density_matrix
    want = O.simulate_density_matrix(to_oracle_ops(nc, qs), 8)
    assert np.linalg.norm(got - want) < 1e-5
    ev = sim.simulate_expectation_values(nc, [generators.sum_z(qs)])[0]
    want_ev = sum(O.expectation_from_density_matrix(want, t, 8) for t in terms_of(generators.sum_z(qs), qs))
    assert abs(ev - want_ev) < 1e-5


@pytest.mark.parametrize("seed", range(3))
def test_density_matrix_random_channels(gpu, seed):
    rng = np.random.RandomState(200 + seed)
    c, qs = _random_circuit(5, 60, rng, with_channels=True)
    got = cirq.DensityMatrixSimulator(tile_bits=6, coalesce_bits=2).simulate(c, qubit_order=qs).final_density_matrix
    want = O.simulate_density_matrix(to_oracle_ops(c, qs), 5)
    assert np.linalg.norm(got - want) < 1e-5


def test_density_matrix_sampling_known_answer(gpu):
    # cirq-core/cirq/sim/density_matrix_utils_test.py:77-93 through the simulator's run():
    # H then measure gives p = (1/2, 1/2) from diag(rho)
    q = cirq.LineQubit(0)
    c = cirq.Circuit(cirq.H(q), cirq.measure(q, key="m"))
    r = cirq.DensityMatrixSimulator(seed=1234).run(c, repetitions=10).measurements["m"][:, 0]
    assert r.tolist() == [0, 1, 0, 1, 1, 0, 0, 1, 1, 1]


def test_parameter_sweep_matches_oracle(gpu):
    circ, syms, vals = generators.hardware_efficient_ansatz(8, 3, seed=1)
    qs = sorted(circ.all_qubits())
    rng = np.random.RandomState(3)
    resolvers = [cirq.ParamResolver(dict(zip(syms, rng.uniform(-3, 3, len(syms))))) for _ in range(4)]
    res = cirq.Simulator().simulate_sweep(circ, resolvers)
    for r, rr in zip(resolvers, res):
        want = O.simulate_state_vector(to_oracle_ops(circ, qs, resolver=r), 8)
        assert np.max(np.abs(rr.final_state_vector - want)) <= ATOL


def test_norm_and_inverse_round_trip_at_26_qubits(gpu):
    """Size-independent checks at a large size: norm preserved, U^-1 U |0> = |0>."""
    _log("26q round trip")
    c = generators.grid_circuit(5, 6, 12, seed=5, n_qubits=26)
    qs = sorted(c.all_qubits())
    inv = cirq.Circuit([cirq.MatrixGate(np.linalg.inv(op.gate._unitary_())).on(*op.qubits) for op in reversed(list(c.all_operations()))])
    sim = cirq.Simulator()
    comp = sim._compile(c, qs)
    comp_inv = sim._compile(inv, qs)
    st = engine.new_state(26, sim._dev())
    comp.dev.run(st)
    n2 = engine.norm2(st, 26, sim._ws)
    assert abs(n2 - 1) < 1e-5
    comp_inv.dev.run(st)
    a0 = st[0].item()
    assert abs(abs(a0) ** 2 - 1) < 1e-5


def test_unsupported_ops_raise_type_error(gpu):
    q = cirq.LineQubit.range(2)

    class Opaque(cirq.Gate):  # no unitary, mixture or Kraus operators
        def _num_qubits_(self):
            return 1

    with pytest.raises(TypeError):
        cirq.Simulator().simulate(cirq.Circuit(Opaque().on(q[0])))
    with pytest.raises(ValueError):
        cirq.Simulator().run(cirq.Circuit(cirq.H(q[0])))
    # a channel on a state vector is sampled as one trajectory (a normalised state)
    psi = cirq.Simulator(seed=1).simulate(cirq.Circuit(cirq.depolarize(0.1).on(q[0]))).final_state_vector
    assert abs(np.linalg.norm(psi) - 1) < 1e-6


@pytest.mark.parametrize("dm", [False, True])
def test_passes_from_basis_state_match_init_then_passes(gpu, dm):
    """cqs_run_passes_from_basis builds |b> in the first pass: bit-identical to cqs_init_basis
    followed by cqs_run_passes, for several tiles and basis states, and with no pass to run."""
    _log(f"passes from a basis state dm={dm}")
    from cirq_amd import engine, tape

    rng = np.random.RandomState(8)
    n = 5 if dm else 11
    c, qs = _random_circuit(n, 60, rng, with_channels=dm)
    low = tape.Lowering(qs, dm)
    low.lower_circuit(c)
    nb = 2 * n if dm else n
    dev = torch.device("cuda")
    for L, co in [(4, 2), (7, 3), (min(nb, 10), 4)]:
        prog = engine.DeviceProgram(tape.compile_ops(low.ops, nb, L=L, coalesce=co), dev)
        for b in (0, 1, 37 % 2**n, 2**n - 1):
            idx = b * 2**n + b if dm else b
            a = engine.new_state(nb, dev, idx)
            prog.run(a)
            x = torch.full((1 << nb,), complex(np.nan, np.nan), dtype=torch.complex64, device=dev)  # never read
            prog.run(x, basis=idx)
            assert torch.equal(a, x), (L, b)
    empty = engine.DeviceProgram(tape.compile_ops([], nb), dev)
    x = torch.full((1 << nb,), 7.0, dtype=torch.complex64, device=dev)
    empty.run(x, basis=3)
    want = torch.zeros(1 << nb, dtype=torch.complex64, device=dev)
    want[3] = 1
    assert torch.equal(x, want)
