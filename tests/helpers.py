"""Test helpers: circuit -> oracle op conversion and a NumPy model of the pass kernel.

``to_oracle_ops`` describes each operation by its gate family and public parameters so that the
oracle computes every matrix from its own tables (not from cirq_amd's gate code).

``run_pass_model`` executes a PackedProgram with the semantics the HIP pass kernel implements
(tile gather over local bits, dense ops on tile bits with local/outer controls, diagonals and pair
phases on global bits, scales).  It lets the host compiler (lowering, fusion, scheduling,
packing) be tested on a machine without a GPU.
"""

from __future__ import annotations

import numpy as np

from cirq_amd import _lib


def _gate_spec(gate):
    name = type(gate).__name__
    if name in ("_PauliX", "_PauliY", "_PauliZ"):
        return "Pauli", {"pauli": name[-1]}
    if name in ("Rx", "Ry", "Rz"):
        return name, {"rads": float(gate.rads)}
    if name in (
        "XPowGate", "YPowGate", "ZPowGate", "HPowGate", "CZPowGate", "CXPowGate", "SwapPowGate",
        "ISwapPowGate", "XXPowGate", "YYPowGate", "ZZPowGate", "CCZPowGate", "CCXPowGate",
    ):
        return name, {"exponent": float(gate.exponent), "global_shift": float(gate.global_shift)}
    if name == "PhasedXPowGate":
        return name, {
            "phase_exponent": float(gate.phase_exponent),
            "exponent": float(gate.exponent),
            "global_shift": float(gate.global_shift),
        }
    if name == "PhasedXZGate":
        return name, {
            "x_exponent": float(gate.x_exponent),
            "z_exponent": float(gate.z_exponent),
            "axis_phase_exponent": float(gate.axis_phase_exponent),
        }
    if name == "FSimGate":
        return name, {"theta": float(gate.theta), "phi": float(gate.phi)}
    if name == "DiagonalGate":
        return name, {"angles": [float(a) for a in gate.diag_angles_radians]}
    if name == "ControlledGate":
        sf, sp = _gate_spec(gate.sub_gate)
        return "Controlled", {"sub_family": sf, "sub_params": sp, "control_values": list(gate.control_values)}
    if name == "IdentityGate":
        return name, {"num_qubits": gate.num_qubits()}
    if name == "CSwapGate":
        return name, {}
    return "MatrixGate", {"matrix": np.asarray(gate._unitary_())}


def _channel_spec(gate):
    name = type(gate).__name__
    if name == "DepolarizingChannel":
        return "depolarize", {"p": gate.p, "n_qubits": gate.n_qubits}
    if name == "AsymmetricDepolarizingChannel":
        return "asymmetric_depolarize", {"error_probabilities": dict(gate.error_probabilities)}
    if name == "AmplitudeDampingChannel":
        return "amplitude_damp", {"gamma": gate.gamma}
    if name == "GeneralizedAmplitudeDampingChannel":
        return "generalized_amplitude_damp", {"p": gate.p, "gamma": gate.gamma}
    if name == "PhaseDampingChannel":
        return "phase_damp", {"gamma": gate.gamma}
    if name == "BitFlipChannel":
        return "bit_flip", {"p": gate.p}
    if name == "PhaseFlipChannel":
        return "phase_flip", {"p": gate.p}
    if name == "ResetChannel":
        return "reset", {}
    return "kraus", {"ops": list(gate._kraus_())}


def to_oracle_ops(circuit, qubits, resolver=None):
    idx = {q: i for i, q in enumerate(qubits)}
    out = []
    for moment in circuit:
        for op in moment:
            if resolver is not None and hasattr(op, "_resolve_parameters_"):
                op = op._resolve_parameters_(resolver, True)
            gate = getattr(op, "gate", None)
            axes = tuple(idx[q] for q in op.qubits)
            if gate is None:  # a multi-qubit PauliString
                for q, p in op.items():
                    out.append(("gate", "Pauli", {"pauli": p._name}, (idx[q],)))
                if op.coefficient != 1:
                    out.append(("gate", "GlobalPhase", {"coefficient": complex(op.coefficient)}, ()))
                continue
            name = type(gate).__name__
            if name == "MeasurementGate":
                continue
            if name == "GlobalPhaseGate":
                out.append(("gate", "GlobalPhase", {"coefficient": complex(gate.coefficient)}, ()))
                continue
            if hasattr(gate, "_kraus_") or hasattr(gate, "_mixture_"):
                fam, params = _channel_spec(gate)
                out.append(("channel", fam, params, axes))
                continue
            fam, params = _gate_spec(gate)
            out.append(("gate", fam, params, axes))
    return out


def to_oracle_moments(circuit, qubits, resolver=None):
    """The circuit as moments of neutral oracle ops, measurements included as
    ('measure', key, {'invert': full invert mask}, axes)."""
    idx = {q: i for i, q in enumerate(qubits)}
    out = []
    for moment in circuit:
        ops = []
        for op in moment:
            gate = getattr(op, "gate", None)
            if gate is not None and type(gate).__name__ == "MeasurementGate":
                ops.append(("measure", gate.key, {"invert": list(gate.full_invert_mask())},
                            tuple(idx[q] for q in op.qubits)))
            else:
                ops += to_oracle_ops([[op]], qubits, resolver)
        out.append(ops)
    return out


def terms_of(psum, qubits):
    """PauliSum -> oracle terms [(coefficient, {axis: 'X'|'Y'|'Z'})]."""
    idx = {q: i for i, q in enumerate(qubits)}
    return [(complex(t.coefficient), {idx[q]: p._name for q, p in t.items()}) for t in psum]


# ------------------------------------------------------------------------------ pass model
def _deposit(x, pos):
    r = np.zeros_like(x, dtype=np.int64)
    for b, p in enumerate(pos):
        r |= ((x >> b) & 1) << p
    return r


def _glob(code: int, lpos) -> int:
    """Global bit of a DIAG / PAIRPHASE position code (tile-local index or 64 + global bit)."""
    return code - 64 if code >= 64 else lpos[code]


def _apply_op(r, tile, gi, base, L, lpos, mats, mats_c64, aux):
    """One non-gradient op of a pass on the gathered tiles (ntiles, 2^L); returns the new tiles."""
    kind = int(r["kind"]) & 0xFF
    com = int(r["co_mask_lo"]) | (int(r["co_mask_hi"]) << 32)
    cov = int(r["co_val_lo"]) | (int(r["co_val_hi"]) << 32)
    tile_ok = (base & com) == cov  # (ntiles,)
    if kind in (_lib.OP_DENSE, _lib.OP_BLOCK):
        k = int(r["k"])
        t = [int(x) for x in r["t"][:k]]
        D = 1 << k
        # matrix variants selected by outer bits (t[5] = count << 24 | 7-bit positions)
        code = int(r["t"][5])
        vpos = [(code >> (7 * i)) & 127 for i in range((code >> 24) & 3)]
        variant = np.zeros_like(base)
        for i, b in enumerate(vpos):
            variant |= ((base >> b) & 1) << i
        stride = D * D if kind == _lib.OP_DENSE else 2 * D * D

        def matrix(v):
            o = int(r["mat"]) + v * stride
            if kind == _lib.OP_DENSE:
                return mats[o : o + D * D].reshape(D, D)
            from cirq_amd.tape import block_matrix_from_payload

            return block_matrix_from_payload(np.asarray(mats_c64[o : o + 2 * D * D], dtype=np.complex64), k)

        j = np.arange(1 << L)
        j0 = j[np.all([(j >> b) & 1 == 0 for b in t], axis=0)]
        pat = np.array([sum(((m >> b) & 1) << t[b] for b in range(k)) for m in range(D)])
        idx = j0[:, None] + pat[None, :]  # (G, D)
        grp_ok = (j0 & int(r["cl_mask"])) == int(r["cl_val"])
        V = tile[:, idx]  # (ntiles, G, D)
        out = np.empty_like(V)
        for v in np.unique(variant):
            sel = variant == v
            out[sel] = V[sel] @ matrix(int(v)).T
        ok = tile_ok[:, None, None] & grp_ok[None, :, None]
        tile = tile.copy()
        tile[:, idx] = np.where(ok, out, V)
        return tile
    if kind == _lib.OP_DIAG:
        k = int(r["k"])
        t = [_glob(int(x), lpos) for x in r["t"][:k]]
        Dv = mats[int(r["mat"]) : int(r["mat"]) + (1 << k)]
        di = np.zeros_like(gi)
        for m, b in enumerate(t):
            di |= ((gi >> b) & 1) << m
        return np.where(tile_ok[:, None], tile * Dv[di], tile)
    if kind == _lib.OP_PAIRPHASE:
        k = int(r["k"])
        a0 = int(r["t"][0])
        ph = mats[int(r["mat"]) : int(r["mat"]) + k]
        for p in range(k):
            a, b = _glob(int(aux[a0 + 2 * p]), lpos), _glob(int(aux[a0 + 2 * p + 1]), lpos)
            both = ((gi >> a) & (gi >> b) & 1).astype(bool)
            tile = np.where(both & tile_ok[:, None], tile * ph[p], tile)
        return tile
    if kind == _lib.OP_SCALE:
        return np.where(tile_ok[:, None], tile * mats[int(r["mat"])], tile)
    raise ValueError(kind)


def _gather(packed, ps):
    n = packed.n_bits
    L = int(ps["L"])
    lpos = [int(x) for x in ps["lpos"][:L]]
    opos = [p for p in range(n) if p not in lpos]
    tiles = np.arange(1 << (n - L), dtype=np.int64)
    base = _deposit(tiles, opos)  # (ntiles,)
    off = _deposit(np.arange(1 << L, dtype=np.int64), lpos)  # (T,)
    return L, lpos, base, base[:, None] + off[None, :]


def _permute(r, tile, I, lpos, aux):
    """PERMUTE: tile bit b moves to tile bit aux[t0 + b] (tiles and their flat indices alike)."""
    L = len(lpos)
    dst = [int(aux[int(r["t"][0]) + b]) for b in range(L)]
    assert sorted(dst) == list(range(L))
    j = np.arange(1 << L)
    nj = np.zeros_like(j)
    for b in range(L):
        nj |= ((j >> b) & 1) << dst[b]
    nt = np.empty_like(tile)
    nI = np.empty_like(I)
    nt[:, nj] = tile
    nI[:, nj] = I
    nl = [0] * L
    for b in range(L):
        nl[dst[b]] = lpos[b]
    return nt, nI, nl


def run_pass_model(packed, state: np.ndarray, mats: np.ndarray | None = None) -> np.ndarray:
    """Executes the packed program on a copy of `state` (flat complex array)."""
    st = np.array(state, dtype=np.complex128)
    mats_c64 = np.asarray(packed.mats_arr if mats is None else mats, dtype=np.complex64)
    mats = mats_c64.astype(np.complex128)
    for ps in packed.pass_arr:
        L, lpos, base, I = _gather(packed, ps)
        tile = st[I]
        for r in packed.ops_arr[int(ps["op_begin"]) : int(ps["op_begin"]) + int(ps["op_count"])]:
            if int(r["kind"]) == _lib.OP_PERMUTE:
                tile, I, lpos = _permute(r, tile, I, lpos, packed.aux_arr)
                continue
            tile = _apply_op(r, tile, I, base, L, lpos, mats, mats_c64, packed.aux_arr)
        assert lpos == [int(x) for x in ps["lpos_out"][:L]], "lpos_out disagrees with the PERMUTE ops"
        st[I] = tile
    return st


def run_adjoint_model(packed, psi: np.ndarray, phi: np.ndarray, mats: np.ndarray, n_slots: int):
    """The reverse sweep of the adjoint kernel: tile selectors (kind bits 8-9: 0 both, 1 psi,
    2 phi) and GRAD ops accumulating 2 Re <phi| dU |psi> per slot.  Returns (psi, phi, grads)."""
    psi = np.array(psi, dtype=np.complex128)
    phi = np.array(phi, dtype=np.complex128)
    mats_c64 = np.asarray(mats, dtype=np.complex64)
    mats = mats_c64.astype(np.complex128)
    grads = np.zeros(n_slots)
    for ps in packed.pass_arr:
        L, lpos, base, I = _gather(packed, ps)
        ta, tb = psi[I], phi[I]
        for r in packed.ops_arr[int(ps["op_begin"]) : int(ps["op_begin"]) + int(ps["op_count"])]:
            kind, sel = int(r["kind"]) & 0xFF, (int(r["kind"]) >> 8) & 3
            if kind in (_lib.OP_GRAD_DENSE, _lib.OP_GRAD_DIAG):
                rr = r.copy()
                rr["kind"] = _lib.OP_DENSE if kind == _lib.OP_GRAD_DENSE else _lib.OP_DIAG
                y = _apply_op(rr, ta, I, base, L, lpos, mats, mats_c64, packed.aux_arr)  # dU psi
                grads[int(r["slot"])] += 2 * np.real(np.vdot(tb, y))
                continue
            if sel != 2:
                ta = _apply_op(r, ta, I, base, L, lpos, mats, mats_c64, packed.aux_arr)
            if sel != 1:
                tb = _apply_op(r, tb, I, base, L, lpos, mats, mats_c64, packed.aux_arr)
        psi[I], phi[I] = ta, tb
    return psi, phi, grads


def fidelity(a: np.ndarray, b: np.ndarray) -> float:
    a = np.asarray(a, dtype=np.complex128).reshape(-1)
    b = np.asarray(b, dtype=np.complex128).reshape(-1)
    return float(abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real))


def _pauli_raw(psi, x, z):
    i = np.arange(psi.size, dtype=np.int64)
    sign = np.where(np.array([bin(v).count("1") & 1 for v in (i & z)]), -1.0, 1.0)
    return np.sum(sign * np.conj(psi[i ^ x]) * psi)


def run_batch_model(h, mats_arena):
    """Executes a planned BatchPrograms (cirq_amd.batch) launch by launch with the pass models:
    returns (raw <P_t> per point and term, flat; gradient slots per point, flat)."""
    from types import SimpleNamespace

    arena = np.zeros(h.arena_elems, dtype=np.complex128)
    point_of = {}
    for i, (c, _) in enumerate(h.points):
        point_of[int(h.term_items[i]["state"])] = i
    grads = {i: np.zeros(h.progs[c].n_slots) for i, (c, _) in enumerate(h.points)}
    raw = None

    def before_adjoint():
        out = []
        for i, ti in enumerate(h.term_items):
            n, st = int(ti["n_bits"]), int(ti["state"])
            psi = arena[st : st + (1 << n)].copy()
            phi = np.zeros_like(psi)
            for t in range(int(ti["term_begin"]), int(ti["term_begin"]) + int(ti["n_terms"])):
                x, z = int(h.xmask[t]), int(h.zmask[t])
                out.append(_pauli_raw(psi, x, z))
                j = np.arange(psi.size, dtype=np.int64)
                src = j ^ x
                sign = np.where(np.array([bin(v).count("1") & 1 for v in (src & z)]), -1.0, 1.0)
                phi += h.apply_coef[t] * sign * psi[src]
            if h.gradients:
                st2 = int(ti["state2"])
                arena[st2 : st2 + (1 << n)] = phi.astype(np.complex64)
        return np.array(out, dtype=np.complex128)

    for kind, L, off, n_items, n_tiles in h.launches:
        if kind == "adj" and raw is None:
            raw = before_adjoint()
        for it in h.items[off : off + n_items]:
            B = h.passes[int(it["pass"])]
            n, st = int(B["n_bits"]), int(it["state"])
            if kind == "init":
                arena[st : st + (1 << n)] = 0
                arena[st] = 1
                continue
            row = np.zeros(1, dtype=_lib.PASS_DTYPE)
            row["L"], row["op_begin"], row["op_count"], row["max_k"] = B["L"], B["op_begin"], B["op_count"], B["max_k"]
            row["lpos"], row["lpos_out"] = B["lpos"], B["lpos_out"]
            view = SimpleNamespace(n_bits=n, pass_arr=row, ops_arr=h.ops, aux_arr=h.aux, mats_arr=None)
            mats = mats_arena[int(it["mats"]) :]
            if kind == "fwd":
                arena[st : st + (1 << n)] = run_pass_model(view, arena[st : st + (1 << n)], mats)
            else:
                st2 = int(it["state2"])
                p = point_of[st]
                psi, phi, g = run_adjoint_model(view, arena[st : st + (1 << n)], arena[st2 : st2 + (1 << n)], mats,
                                                len(grads[p]))
                arena[st : st + (1 << n)], arena[st2 : st2 + (1 << n)] = psi, phi
                grads[p] += g
    if raw is None:
        raw = before_adjoint()
    slots = np.concatenate([grads[i] for i in range(len(h.points))]) if h.gradients else None
    return raw, slots


def vec_observable(row_terms, n_bits: int) -> np.ndarray:
    """vec(H) as a 2n-bit density-matrix vector: the Pauli sum (masks on the row bits) applied to
    vec(I), with the same phase convention as the expectation kernels."""
    n = n_bits // 2
    eye = np.zeros(1 << n_bits, dtype=np.complex128)
    eye[np.arange(1 << n) * ((1 << n) + 1)] = 1
    j = np.arange(1 << n_bits, dtype=np.int64)
    out = np.zeros_like(eye)
    for x, z, c in row_terms:
        src = j ^ x
        sign = np.where(np.array([bin(v).count("1") & 1 for v in (src & z)]), -1.0, 1.0)
        out += c * sign * eye[src]
    return out


def run_dm_adjoint_model(prog, values):
    """DMAdjointGradient.value_and_grad step by step on the pass models: the forward pass with
    checkpoints, vec(H), then every segment recomputed from its checkpoint and walked backwards."""
    from types import SimpleNamespace

    symvals = prog.symbol_values(values)
    mf, mr = prog.mats(symvals)
    nb = prog.n_bits

    def sub(packed, a, b):
        return SimpleNamespace(n_bits=nb, pass_arr=packed.pass_arr[a:b], ops_arr=packed.ops_arr,
                               aux_arr=packed.aux_arr, mats_arr=None)

    def step(j, st):
        a, b = prog.fwd_range[j]
        return run_pass_model(sub(prog.fwd, a, b), st, mf) if b > a else st

    rho0 = np.zeros(1 << nb, dtype=np.complex128)
    rho0[0] = 1
    segs = prog.segments()
    ckpt = []
    rho = rho0.copy()
    for g, (a, b) in enumerate(segs):
        if g > 0:
            ckpt.append(rho.copy())
        for j in range(a, b):
            rho = step(j, rho)
    lam = vec_observable(prog.row_terms, nb)  # vec(H) / phi_scale
    value = float(np.vdot(lam, rho).real) * prog.phi_scale
    grads = np.zeros(prog.n_slots)
    for g in range(len(segs) - 1, -1, -1):
        a, b = segs[g]
        rho = rho0.copy() if g == 0 else ckpt[g - 1].copy()
        stretch = []
        for j in range(a, b):
            rho = step(j, rho)
            stretch.append(rho)
        for j in range(b - 1, a - 1, -1):
            ra, rb = prog.rev_range[j]
            if rb > ra:
                _, lam, gj = run_adjoint_model(sub(prog.rev, ra, rb), stretch[j - a], lam, mr, prog.n_slots)
                grads += gj
    return value, prog.chain(grads, symvals), rho
