"""GPU tests of moment steps, mid-circuit measurements and noisy state-vector trajectories (the
behaviour of cirq-core/cirq/sim/sparse_simulator_test.py:215-300,566-630 restated)."""

import sys

import numpy as np
import pytest

import cirq_amd as cirq
from oracle import cirq_np as O
from tests.helpers import to_oracle_ops

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_moment_steps_states(gpu):
    _log("moment steps")
    q0, q1 = cirq.LineQubit.range(2)
    circuit = cirq.Circuit(cirq.H(q0), cirq.H(q1), cirq.H(q0), cirq.H(q1))
    steps = list(cirq.Simulator().simulate_moment_steps(circuit))
    assert len(steps) == 2
    np.testing.assert_allclose(steps[0].state_vector(), [0.5] * 4, atol=1e-6)
    np.testing.assert_allclose(steps[1].state_vector(), [1, 0, 0, 0], atol=1e-6)
    steps = list(cirq.Simulator().simulate_moment_steps(cirq.Circuit()))
    assert len(steps) == 1 and np.allclose(steps[0].state_vector(), [1]) and not steps[0].qubit_map


def test_moment_steps_match_oracle_prefixes(gpu):
    _log("moment steps vs oracle prefixes")
    from tests.test_host import _random_circuit

    c, qs = _random_circuit(7, 12, np.random.RandomState(3))
    sim = cirq.Simulator()
    moments = list(c)
    for i, step in enumerate(sim.simulate_moment_steps(c, qubit_order=qs)):
        want = O.simulate_state_vector(to_oracle_ops(cirq.Circuit.from_moments(*moments[: i + 1]), qs), 7)
        assert np.max(np.abs(step.state_vector() - want)) < 1e-5


def test_moment_steps_sample_and_intermediate_measurement(gpu):
    _log("moment steps sample / intermediate measurement")
    q0, q1 = cirq.LineQubit.range(2)
    for i, step in enumerate(cirq.Simulator().simulate_moment_steps(cirq.Circuit(cirq.H(q0), cirq.CNOT(q0, q1)))):
        s = step.sample([q0, q1], repetitions=10)
        allowed = ([1, 0], [0, 0]) if i == 0 else ([1, 1], [0, 0])
        assert all(list(x) in [list(a) for a in allowed] for x in s)
    for seed in range(4):
        circuit = cirq.Circuit(cirq.H(q0), cirq.measure(q0), cirq.H(q0))
        for i, step in enumerate(cirq.Simulator(seed=seed).simulate_moment_steps(circuit)):
            if i == 1:
                result = int(step.measurements["q(0)"][0])
                expected = np.zeros(2)
                expected[result] = 1
                np.testing.assert_allclose(step.state_vector(), expected, atol=1e-6)
            if i == 2:
                np.testing.assert_allclose(step.state_vector(), [np.sqrt(0.5), np.sqrt(0.5) * (-1) ** result],
                                           atol=1e-6)


def test_run_measurement_not_terminal(gpu):
    _log("run with non-terminal measurements")
    q0, q1 = cirq.LineQubit.range(2)
    sim = cirq.Simulator()
    for b0 in (0, 1):
        for b1 in (0, 1):
            circuit = cirq.Circuit((cirq.X**b0)(q0), (cirq.X**b1)(q1), cirq.measure(q0), cirq.measure(q1),
                                   cirq.H(q0), cirq.H(q1))
            result = sim.run(circuit, repetitions=3)
            np.testing.assert_equal(result.measurements["q(0)"], [[b0]] * 3)
            np.testing.assert_equal(result.measurements["q(1)"], [[b1]] * 3)


def test_mid_circuit_measurement_correlations_and_repeated_keys(gpu):
    _log("mid-circuit correlations")
    q0, q1 = cirq.LineQubit.range(2)
    circuit = cirq.Circuit(cirq.H(q0), cirq.measure(q0, key="a"), cirq.CNOT(q0, q1), cirq.measure(q1, key="b"),
                           cirq.H(q0), cirq.measure(q0, key="c"), cirq.measure(q0, key="c"))
    r = cirq.Simulator(seed=5).run(circuit, repetitions=400)
    a, b, c = r.records["a"], r.records["b"], r.records["c"]
    assert a.shape == (400, 1, 1) and c.shape == (400, 2, 1)
    assert np.array_equal(a[:, 0, 0], b[:, 0, 0])
    assert np.array_equal(c[:, 0, 0], c[:, 1, 0])  # a repeated measurement repeats its outcome
    assert 120 < a.sum() < 280 and 120 < c[:, 0, 0].sum() < 280


def test_state_vector_noise_trajectories(gpu):
    _log("noisy trajectories")
    q0 = cirq.LineQubit(0)
    r = cirq.Simulator(seed=1).run(cirq.Circuit(cirq.bit_flip(0.5)(q0), cirq.measure(q0)), repetitions=100)
    assert 20 < np.count_nonzero(r.measurements["q(0)"]) < 80
    r = cirq.Simulator(seed=23).run(cirq.Circuit(cirq.H(q0), cirq.phase_flip(0.5)(q0), cirq.H(q0), cirq.measure(q0)),
                                    repetitions=100)
    assert 20 < np.count_nonzero(r.measurements["q(0)"]) < 80
    # amplitude damping (Kraus, not a mixture): from |1>, P(1) = 1 - gamma
    circuit = cirq.Circuit(cirq.X(q0), cirq.amplitude_damp(0.3)(q0), cirq.measure(q0))
    r = cirq.Simulator(seed=2).run(circuit, repetitions=2000)
    frac = np.mean(r.measurements["q(0)"])
    assert abs(frac - 0.7) < 0.04
    # the trajectory average of the final state's populations matches the density matrix
    c2 = cirq.Circuit(cirq.H(q0), cirq.depolarize(0.3)(q0), cirq.rx(0.4)(q0))
    rho = cirq.DensityMatrixSimulator().simulate(c2).final_density_matrix
    sim = cirq.Simulator(seed=3)
    pops = np.mean([np.abs(sim.simulate(c2).final_state_vector) ** 2 for _ in range(400)], axis=0)
    assert np.max(np.abs(pops - np.real(np.diag(rho)))) < 0.06


def test_density_matrix_mid_circuit_measurement(gpu):
    _log("density matrix mid-circuit measurement")
    q0, q1 = cirq.LineQubit.range(2)
    circuit = cirq.Circuit(cirq.H(q0), cirq.measure(q0, key="m"), cirq.CNOT(q0, q1), cirq.amplitude_damp(0.1)(q1))
    res = cirq.DensityMatrixSimulator(seed=4).simulate(circuit)
    m = int(res.measurements["m"][0])
    rho = res.final_density_matrix
    assert abs(np.trace(rho) - 1) < 1e-5
    # q0 is now classical: populations only where q0 == m
    diag = np.real(np.diag(rho)).reshape(2, 2)
    assert np.allclose(diag[1 - m], 0, atol=1e-6)


# ---------------------------------------------------------------- batched trajectories
def _agreement(ra, rb, key):
    a, b = ra.records[key], rb.records[key]
    assert a.shape == b.shape
    return float(np.mean(np.all(a.reshape(len(a), -1) == b.reshape(len(b), -1), axis=1)))


def _mid_circuit(n=8, seed=0):
    rng = np.random.RandomState(seed)
    qs = cirq.LineQubit.range(n)
    c = cirq.Circuit()
    for layer in range(4):
        for q in qs:
            c.append(cirq.rx(rng.uniform(-3, 3))(q))
            c.append(cirq.rz(rng.uniform(-3, 3))(q))
        for i in range(layer % 2, n - 1, 2):
            c.append(cirq.CNOT(qs[i], qs[i + 1]))
        if layer == 1:
            c.append(cirq.measure(qs[1], qs[4], key="mid"))
        if layer == 2:
            c.append(cirq.measure(qs[6], key="again"))
            c.append(cirq.measure(qs[6], key="again"))
    c.append(cirq.measure(*qs, key="end"))
    return c


def test_batched_trajectories_match_one_at_a_time(gpu):
    """Every repetition consumes the same uniforms in both paths; outcomes agree except where a
    probability rounds differently at a draw's boundary."""
    _log("batched vs one-at-a-time trajectories (mid-circuit measurements)")
    c = _mid_circuit()
    sb = cirq.Simulator(seed=11)
    ss = cirq.Simulator(seed=11, trajectory_batch=1)
    rb = sb.run(c, repetitions=300)
    rs = ss.run(c, repetitions=300)
    for key in ("mid", "again", "end"):
        assert _agreement(rb, rs, key) >= 0.99, key
    assert rb.records["again"].shape == (300, 2, 1)
    assert np.array_equal(rb.records["again"][:, 0], rb.records["again"][:, 1])
    assert sb._prng.random_sample() == ss._prng.random_sample()  # the same draws were consumed


def test_batched_trajectory_chunks(gpu):
    """Batches of 64 copies (the last one padded) give the records of one batch of 1024."""
    _log("batched trajectories in chunks")
    c = _mid_circuit(seed=2)
    r1 = cirq.Simulator(seed=5).run(c, repetitions=1000)
    r2 = cirq.Simulator(seed=5, trajectory_batch=64).run(c, repetitions=1000)
    for key in ("mid", "again", "end"):
        assert _agreement(r1, r2, key) >= 0.995, key


def _noisy_circuit():
    qs = cirq.LineQubit.range(6)
    c = cirq.Circuit(cirq.H.on_each(*qs))
    c.append(cirq.CNOT(qs[i], qs[i + 1]) for i in range(5))
    c.append([cirq.depolarize(0.08)(qs[0]), cirq.depolarize(0.05, n_qubits=2)(qs[1], qs[2]),
              cirq.amplitude_damp(0.3)(qs[3]), cirq.bit_flip(0.1)(qs[4]), cirq.phase_damp(0.2)(qs[5])])
    c.append(cirq.rx(0.7)(q) for q in qs)
    c.append([cirq.amplitude_damp(0.25)(qs[0]), cirq.generalized_amplitude_damp(0.3, 0.2)(qs[2])])
    c.append(cirq.ry(0.4)(q) for q in qs)
    return c, qs


def test_batched_noisy_trajectories(gpu):
    """Mixtures and Kraus channels on a state vector: the batched path agrees with the
    one-at-a-time path and its outcome frequencies match the density-matrix probabilities."""
    _log("batched noisy trajectories")
    c, qs = _noisy_circuit()
    meas = c + cirq.Circuit(cirq.measure(qs[0], qs[2], qs[3], key="m"))
    rb = cirq.Simulator(seed=3).run(meas, repetitions=300)
    rs = cirq.Simulator(seed=3, trajectory_batch=1).run(meas, repetitions=300)
    assert _agreement(rb, rs, "m") >= 0.98
    rho = cirq.DensityMatrixSimulator().simulate(c, qubit_order=qs).final_density_matrix
    p = np.real(np.diag(rho)).reshape([2] * 6).transpose(0, 2, 3, 1, 4, 5).reshape(8, 8).sum(axis=1)
    big = cirq.Simulator(seed=9).run(meas, repetitions=6000).records["m"][:, 0, :]
    idx = big[:, 0] * 4 + big[:, 1] * 2 + big[:, 2]
    freq = np.bincount(idx, minlength=8) / len(idx)
    assert 0.5 * np.abs(freq - p).sum() < 0.035, (freq, p)


def test_batched_density_matrix_mid_circuit(gpu):
    _log("batched density-matrix trajectories")
    q0, q1, q2 = cirq.LineQubit.range(3)
    c = cirq.Circuit(cirq.H(q0), cirq.CNOT(q0, q1), cirq.amplitude_damp(0.2)(q1), cirq.measure(q1, key="a"),
                     cirq.rx(0.9)(q2), cirq.CNOT(q1, q2), cirq.depolarize(0.1)(q2), cirq.measure(q0, q2, key="b"))
    rb = cirq.DensityMatrixSimulator(seed=4).run(c, repetitions=400)
    rs = cirq.DensityMatrixSimulator(seed=4, trajectory_batch=1).run(c, repetitions=400)
    for key in ("a", "b"):
        assert _agreement(rb, rs, key) >= 0.99, key
    a = rb.records["a"][:, 0, 0]
    assert 0.3 < a.mean() < 0.5  # P(q1 = 1) = 0.5 * (1 - 0.2)


def test_batched_trajectories_parameterised_sweep_and_noise_model(gpu):
    """Symbols resolved per sweep point in the batched programs; a noise model's channels sampled
    per repetition.  Both agree with the one-at-a-time path."""
    _log("batched trajectories: parameter sweep and noise model")
    import sympy

    t = sympy.Symbol("t")
    qs = cirq.LineQubit.range(5)
    c = cirq.Circuit(cirq.rx(t)(q) for q in qs)
    c.append(cirq.CNOT(qs[i], qs[i + 1]) for i in range(4))
    c.append(cirq.measure(qs[0], qs[2], key="mid"))
    c.append(cirq.ry(2 * t)(q) for q in qs)
    c.append(cirq.measure(*qs, key="end"))
    sweep = cirq.Points("t", [0.3, 1.7])
    rb = cirq.Simulator(seed=21).run_sweep(c, sweep, repetitions=200)
    rs = cirq.Simulator(seed=21, trajectory_batch=1).run_sweep(c, sweep, repetitions=200)
    for a, b in zip(rb, rs):
        for key in ("mid", "end"):
            assert _agreement(a, b, key) >= 0.99, key
    noisy = cirq.Circuit(cirq.H.on_each(*qs), cirq.CNOT(qs[0], qs[1]), cirq.measure(*qs, key="m"))
    nb = cirq.Simulator(seed=2, noise=cirq.depolarize(0.05)).run(noisy, repetitions=300)
    ns = cirq.Simulator(seed=2, noise=cirq.depolarize(0.05), trajectory_batch=1).run(noisy, repetitions=300)
    assert _agreement(nb, ns, "m") >= 0.98


@pytest.mark.parametrize("m", [1, 3, 13, 14, 16])
def test_draw_rows_bit_exact_with_numpy_choice(gpu, m):
    """cqs_draw_rows: per row, clip, divide by NumPy's float32 sum (one pairwise sum per 8192
    elements, added in order), then RandomState.choice with the row's uniform."""
    _log(f"draw rows m={m}")
    import torch

    from cirq_amd import engine

    rng = np.random.RandomState(m)
    rows = 40 if m < 16 else 12
    q = (rng.exponential(size=(rows, 2**m)) * rng.uniform(0.1, 3.0, size=(rows, 1))).astype(np.float32)
    q[rng.uniform(size=q.shape) < 0.1] = 0
    q[rng.uniform(size=q.shape) < 0.01] = -1e-9  # clipped to zero, as np.clip does
    q[:, -1] = np.abs(q[:, -1]) + np.float32(1e-3)  # no all-zero row
    u = np.random.RandomState(100 + m).random_sample(rows)
    ws = engine.Workspace(torch.device("cuda"))
    idx, p = engine.draw_rows(torch.from_numpy(q).cuda(), u, ws)
    a = np.random.RandomState(100 + m)
    for r in range(rows):
        pr = np.clip(q[r], 0, None)
        pr /= pr.sum()
        want = int(a.choice(len(pr), p=pr))
        assert idx[r] == want and p[r] == pr[want], (r, idx[r], want)


def test_kraus_choose_on_device_matches_the_walk(gpu):
    """cqs_kraus_choose_batch picks what the host restatement of Cirq's Kraus walk picks, with the
    same 1/sqrt(weight) scale."""
    _log("kraus choice on the device")
    import torch

    from cirq_amd import engine
    from cirq_amd.trajectories import kraus_draws

    rng = np.random.RandomState(3)
    R, M = 4000, 6
    W = rng.exponential(size=(R, M))
    W[rng.uniform(size=W.shape) < 0.25] = 0.0
    W /= W.sum(axis=1, keepdims=True) + rng.uniform(-0.05, 0.05, size=(R, 1))
    W[:3, 1:] = 0.0  # walks that end on the first operator or fall back to it
    u = rng.uniform(size=R)
    sel, scl = engine.kraus_choose_batch(torch.from_numpy(W).cuda(), torch.from_numpy(u).cuda())
    want_sel, want_w = kraus_draws(W, u)
    assert np.array_equal(sel.cpu().numpy(), want_sel)
    with np.errstate(divide="ignore"):
        want_scale = (1.0 / np.sqrt(want_w)).astype(np.float32)
    assert np.array_equal(scl.cpu().numpy(), want_scale)


def test_select_run_matches_one_select_per_op(gpu):
    """cqs_select_run_batch applies a run of per-copy choices (1- and 2-qubit, overlapping
    targets) as the one-op-at-a-time select kernel does, within complex64 rounding."""
    _log("select run vs one select per op")
    import torch

    from cirq_amd import engine

    rng = np.random.RandomState(12)
    n, R = 9, 64
    dev = torch.device("cuda")
    st = torch.from_numpy((rng.randn(R << n) + 1j * rng.randn(R << n)).astype(np.complex64)).to(dev)
    st /= torch.linalg.vector_norm(st.view(R, -1), dim=1).repeat_interleave(1 << n)
    runs = [((2,), 3), ((5,), 4), ((2, 7), 2), ((7,), 3), ((5, 2), 5), ((0,), 2)]  # targets span bits 0,2,5,7
    mats, offs, rows, total = [], [], [], 0
    table = np.zeros((len(runs), R), dtype=np.int32)
    for i, (targets, cnt) in enumerate(runs):
        d = 1 << len(targets)
        ms = np.stack([np.linalg.qr(rng.randn(d, d) + 1j * rng.randn(d, d))[0] for _ in range(cnt)])
        mats.append(ms.astype(np.complex64).reshape(-1))
        offs.append(total)
        total += cnt * d * d
        table[i] = rng.randint(0, cnt, size=R)
    mats_dev = torch.from_numpy(np.concatenate(mats).view(np.float32)).to(dev)
    table_dev = torch.from_numpy(table).to(dev)
    ops = [(len(t), t, offs[i], c, i) for i, (t, c) in enumerate(runs)]
    a = st.clone()
    engine.select_run_batch(a, n, R, ops, mats_dev, table_dev)
    b = st.clone()
    ones = torch.ones(R, dtype=torch.float32, device=dev)
    for i, (targets, cnt) in enumerate(runs):
        d = 1 << len(targets)
        m_dev, _ = engine.device_mats(np.concatenate(mats)[offs[i]:offs[i] + cnt * d * d].reshape(cnt, d, d),
                                      targets, dev)
        engine.apply_select_batch_dev(b, n, R, targets, m_dev, cnt, table_dev[i], ones)
    assert torch.max(torch.abs(a - b)).item() < 2e-6


@pytest.mark.parametrize("same_bit", [False, True])
def test_select_then_kraus_weights_matches_two_kernels(gpu, same_bit):
    """The fused sweep leaves the copies as cqs_apply_select_batch does (bit for bit) and returns
    cqs_kraus_weights_batch's weights up to FP64 summation order."""
    _log(f"select then kraus weights same_bit={same_bit}")
    import torch

    import cirq_amd as cirq
    from cirq_amd import engine
    from cirq_amd.tape import kraus_of

    rng = np.random.RandomState(5)
    n, R = 10, 32
    dev = torch.device("cuda")
    st = torch.from_numpy((rng.randn(R << n) + 1j * rng.randn(R << n)).astype(np.complex64)).to(dev)
    ka = np.stack(kraus_of(cirq.amplitude_damp(0.3)))
    kb = np.stack(kraus_of(cirq.generalized_amplitude_damp(0.4, 0.2)))
    a_bit, b_bit = 3, (3 if same_bit else 7)
    ma, na = engine.device_mats(ka, [a_bit], dev)
    mb, nb = engine.device_mats(kb, [b_bit], dev)
    sel = torch.from_numpy(rng.randint(0, na, size=R).astype(np.int32)).to(dev)
    scl = torch.from_numpy(rng.uniform(0.5, 2.0, size=R).astype(np.float32)).to(dev)
    ws = engine.Workspace(dev)
    x = st.clone()
    w_fused = engine.select_then_kraus_weights(x, n, R, a_bit, ma, na, sel, scl, b_bit, mb, nb, ws).cpu().numpy()
    y = st.clone()
    engine.apply_select_batch_dev(y, n, R, [a_bit], ma, na, sel, scl)
    w_two = engine.kraus_weights_batch_dev(y, n, R, [b_bit], mb, nb, ws)
    assert torch.equal(x, y)
    assert np.max(np.abs(w_fused - w_two) / np.maximum(np.abs(w_two), 1e-300)) < 1e-12


@pytest.mark.parametrize("batch", [None, 1])
def test_trajectories_match_oracle_loop(gpu, batch):
    """Records of run() with a mid-circuit measurement and mixtures / Kraus channels on a state
    vector against the oracle's restatement of Cirq's per-repetition loop (complex64, merged state,
    the same RandomState): every key's record equal in >= 99.9% of repetitions (a draw within
    rounding of a probability boundary may differ), batched and one repetition at a time."""
    from tests.helpers import to_oracle_moments

    q = cirq.LineQubit.range(5)
    c = cirq.Circuit(
        cirq.H.on_each(*q), cirq.CNOT(q[0], q[1]), cirq.ry(0.7).on(q[2]), cirq.measure(q[0], key="a"),
        cirq.depolarize(0.1).on(q[1]), cirq.amplitude_damp(0.25).on(q[2]), cirq.CZ(q[1], q[2]), cirq.X(q[0]) ** 0.3,
        cirq.phase_damp(0.2).on(q[3]), cirq.asymmetric_depolarize(0.05, 0.1, 0.02).on(q[4]), cirq.CNOT(q[3], q[4]),
        cirq.measure(q[1], q[0], key="c", invert_mask=(True,)), cirq.ry(0.3).on(q[0]), cirq.measure(*q, key="b"),
    )
    reps = 2000 if batch is None else 300
    _log(f"trajectories vs oracle loop, batch={batch}, {reps} repetitions")
    got = cirq.Simulator(seed=11, trajectory_batch=batch).run(c, repetitions=reps).records
    want = O.run_trajectories(to_oracle_moments(c, q), 5, reps, np.random.RandomState(11))
    assert set(got) == set(want)
    same = np.ones(reps, dtype=bool)
    for k in want:
        assert got[k].shape == want[k].shape, k
        same &= np.all(got[k] == want[k], axis=(1, 2))
    _log(f"  {int(same.sum())} of {reps} repetitions identical")
    assert same.mean() >= 0.999
