"""Device buffers and C-ABI calls for one state (state vector or density matrix) on one GPU.

All work is enqueued on torch's current HIP stream; only the methods that return NumPy data
synchronise.  There is no CPU fallback: without the built library or a GPU these functions raise.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check, lib, ptr
from .tape import PackedProgram


def require_gpu() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("cirq_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    lib()  # raises if the HIP library is not built
    return torch.device("cuda", torch.cuda.current_device())


def stream_handle() -> int:
    return torch.cuda.current_stream().cuda_stream


def validate_program(packed: PackedProgram, adjoint: bool = False) -> None:
    """cqs_program_validate on the host arrays (ValueError naming the first bad pass / op)."""
    pa = np.ascontiguousarray(packed.pass_arr)
    ops = np.ascontiguousarray(packed.ops_arr)
    aux = np.ascontiguousarray(packed.aux_arr, dtype=np.int32)
    check(lib().cqs_program_validate(int(packed.n_bits), pa.ctypes.data, int(packed.n_passes), ops.ctypes.data,
                                     int(len(ops)), int(packed.mats_arr.size), aux.ctypes.data, int(aux.size),
                                     int(bool(adjoint))), "cqs_program_validate")


class DeviceProgram:
    """A packed pass program resident on the device; parameter matrices can be swapped.  The
    program is checked on the host before upload (validate_program)."""

    def __init__(self, packed: PackedProgram, device, adjoint: bool = False):
        self.packed = packed
        self.n_bits = packed.n_bits
        self.device = device
        validate_program(packed, adjoint)
        self.ops = torch.from_numpy(packed.ops_arr.view(np.int32).copy()).to(device)
        self.mats = torch.from_numpy(packed.mats_arr.view(np.float32).copy()).to(device)
        self.aux = torch.from_numpy(packed.aux_arr.copy()).to(device)
        self.pass_arr = np.ascontiguousarray(packed.pass_arr)
        self.specialized = None  # SpecializedPasses: run() launches them instead of the interpreter

    def specialize(self):
        """Compiles this program's passes to straight-line kernels (cirq_amd.specialize) and runs
        them from now on."""
        from .specialize import specialize

        self.specialized = specialize(self.packed)
        return self

    def set_mats(self, mats_c64: np.ndarray):
        src = torch.from_numpy(np.ascontiguousarray(mats_c64, dtype=np.complex64).view(np.float32))
        self.mats.copy_(src, non_blocking=False)

    def set_mats_device(self, mats_f32: torch.Tensor):
        self.mats = mats_f32

    def run(self, state: torch.Tensor, basis: int | None = None):
        """Runs the passes on state in place; with basis given, on |basis> instead (the first
        pass builds it, so state's contents on entry do not matter)."""
        if self.specialized is not None:
            check(
                lib().cqs_run_passes_specialized(
                    ptr(state), self.n_bits, -1 if basis is None else int(basis), self.pass_arr.ctypes.data,
                    int(self.packed.n_passes), ptr(self.ops), ptr(self.mats), ptr(self.aux),
                    ctypes.cast(self.specialized.handles, ctypes.c_void_p), stream_handle(),
                ),
                "cqs_run_passes_specialized",
            )
            return
        if basis is not None:
            check(
                lib().cqs_run_passes_from_basis(
                    ptr(state), self.n_bits, int(basis), self.pass_arr.ctypes.data, int(self.packed.n_passes),
                    ptr(self.ops), ptr(self.mats), ptr(self.aux), stream_handle(),
                ),
                "cqs_run_passes_from_basis",
            )
            return
        if self.packed.n_passes == 0:
            return
        check(
            lib().cqs_run_passes(
                ptr(state),
                self.n_bits,
                self.pass_arr.ctypes.data,
                int(self.packed.n_passes),
                ptr(self.ops),
                ptr(self.mats),
                ptr(self.aux),
                stream_handle(),
            ),
            "cqs_run_passes",
        )


def check_state_fits(n_bits: int, device) -> None:
    """MemoryError ("Unable to allocate", NumPy's wording, which the reference's simulators
    raise) when 2^n_bits complex64 amplitudes exceed the device's memory."""
    nbytes = 8 << n_bits
    total = torch.cuda.get_device_properties(device).total_memory
    if nbytes > total:
        raise MemoryError(f"Unable to allocate {nbytes / 2**30:.3g} GiB for a state of 2^{n_bits} complex64 "
                          f"amplitudes (the device holds {total / 2**30:.3g} GiB)")


def empty_state(n_bits: int, device) -> torch.Tensor:
    """Uninitialised storage for 2^n_bits complex64 amplitudes; MemoryError when the device
    cannot hold it."""
    nbytes = 8 << n_bits
    check_state_fits(n_bits, device)
    try:
        return torch.empty(1 << n_bits, dtype=torch.complex64, device=device)
    except torch.OutOfMemoryError as e:
        raise MemoryError(f"Unable to allocate {nbytes / 2**30:.3g} GiB for a state of 2^{n_bits} complex64 "
                          "amplitudes") from e


def new_state(n_bits: int, device, basis: int = 0) -> torch.Tensor:
    st = empty_state(n_bits, device)
    check(lib().cqs_init_basis(ptr(st), n_bits, int(basis), stream_handle()), "cqs_init_basis")
    return st


class Workspace:
    """Grow-only device scratch shared by reductions and sampling."""

    def __init__(self, device):
        self.device = device
        self._buf = None

    def get(self, nbytes: int) -> torch.Tensor:
        nbytes = max(256, int(nbytes))
        if self._buf is None or self._buf.numel() < nbytes:
            self._buf = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        return self._buf


def norm2(state: torch.Tensor, n_bits: int, ws: Workspace) -> float:
    out = torch.zeros(2, dtype=torch.float64, device=state.device)
    w = ws.get(lib().cqs_expect_workspace_bytes(1))
    check(lib().cqs_norm2(ptr(state), n_bits, ptr(out), ptr(w), stream_handle()), "cqs_norm2")
    return float(out[0].item())


def inner(a: torch.Tensor, b: torch.Tensor, n_bits: int, ws: Workspace, out: torch.Tensor | None = None):
    """Queues <a|b> over 2^n_bits entries (FP64 accumulation) into out[0:2] (re, im; a new device
    float64 tensor when out is None) and returns out."""
    if out is None:
        out = torch.zeros(2, dtype=torch.float64, device=a.device)
    if out.dtype != torch.float64 or out.numel() < 2 or not out.is_contiguous():
        raise ValueError("out must be a contiguous float64 tensor of at least 2 entries")
    w = ws.get(lib().cqs_expect_workspace_bytes(1))
    check(lib().cqs_inner(ptr(a), ptr(b), n_bits, ptr(out), ptr(w), stream_handle()), "cqs_inner")
    return out


def scale(state: torch.Tensor, n_bits: int, c: complex):
    check(lib().cqs_scale(ptr(state), n_bits, float(c.real), float(c.imag), stream_handle()), "cqs_scale")


def expect_terms(state: torch.Tensor, n_qubits: int, is_dm: bool, terms, ws: Workspace) -> torch.Tensor:
    """terms: list of (xmask, zmask, complex coef incl. i^nY).  Returns device complex128 [T]."""
    T = len(terms)
    out = torch.zeros(2 * max(T, 1), dtype=torch.float64, device=state.device)
    if T == 0:
        return out.view(torch.complex128)[:0]
    xm = np.array([t[0] for t in terms], dtype=np.uint64)
    zm = np.array([t[1] for t in terms], dtype=np.uint64)
    cf = np.array([[complex(t[2]).real, complex(t[2]).imag] for t in terms], dtype=np.float64).reshape(-1)
    w = ws.get(lib().cqs_expect_workspace_bytes(T))
    check(
        lib().cqs_expect_paulis(
            ptr(state), n_qubits, int(is_dm), T, xm.ctypes.data, zm.ctypes.data, cf.ctypes.data, ptr(out), ptr(w),
            stream_handle(),
        ),
        "cqs_expect_paulis",
    )
    return out.view(torch.complex128)


def apply_paulisum(psi: torch.Tensor, phi: torch.Tensor, n_qubits: int, terms):
    T = len(terms)
    xm = np.array([t[0] for t in terms] or [0], dtype=np.uint64)
    zm = np.array([t[1] for t in terms] or [0], dtype=np.uint64)
    cf = np.array([[complex(t[2]).real, complex(t[2]).imag] for t in terms] or [[0, 0]], dtype=np.float64)
    cf = cf.reshape(-1)
    check(
        lib().cqs_apply_paulisum(
            ptr(psi), ptr(phi), n_qubits, T, xm.ctypes.data, zm.ctypes.data, cf.ctypes.data, stream_handle()
        ),
        "cqs_apply_paulisum",
    )


def collapse(state: torch.Tensor, n_bits: int, mask: int, value: int, s: float):
    check(lib().cqs_collapse(ptr(state), n_bits, int(mask), int(value), float(s), 0.0, stream_handle()), "cqs_collapse")


# ---------------------------------------------------------------- batched trajectories
# `states` holds n_states copies of one state one after the other (copy r at r * 2^n_bits).


def _dev_array(x: np.ndarray, device) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(x)).to(device)


def _check_copies(states: torch.Tensor, n_bits: int, n_states: int):
    if states.dtype != torch.complex64 or not states.is_contiguous() or states.numel() < (n_states << n_bits):
        raise ValueError(f"need {n_states} contiguous complex64 copies of 2^{n_bits} amplitudes")


def marginal_batch(states: torch.Tensor, n_qubits: int, is_dm: bool, n_states: int, axes) -> torch.Tensor:
    """Float32 marginal probabilities over `axes` of every copy (device [n_states, 2^m])."""
    _check_copies(states, 2 * n_qubits if is_dm else n_qubits, n_states)
    m = len(axes)
    out = torch.empty((n_states, 1 << m), dtype=torch.float32, device=states.device)
    ax = np.array(axes, dtype=np.int32)
    check(
        lib().cqs_marginal_probs_batch(ptr(states), n_qubits, int(is_dm), n_states, ax.ctypes.data, m, ptr(out),
                                       stream_handle()),
        "cqs_marginal_probs_batch",
    )
    return out


def draw_rows(probs: torch.Tensor, uniforms: np.ndarray, ws: Workspace) -> tuple[np.ndarray, np.ndarray]:
    """NumPy's choice for every row of a device float32 [rows, 2^m] marginal, one uniform per row:
    (outcome index int64[rows], its normalised float32 probability[rows])."""
    rows, width = probs.shape
    m = int(width).bit_length() - 1
    if probs.dtype != torch.float32 or not probs.is_contiguous() or width != 1 << m or len(uniforms) != rows:
        raise ValueError("need a contiguous float32 [rows, 2^m] marginal and one uniform per row")
    u = _dev_array(np.asarray(uniforms, dtype=np.float64), probs.device)
    idx = torch.empty(rows, dtype=torch.int64, device=probs.device)
    p = torch.empty(rows, dtype=torch.float32, device=probs.device)
    w = ws.get(lib().cqs_draw_rows_workspace_bytes(m, rows))
    check(lib().cqs_draw_rows(ptr(probs), m, rows, ptr(u), ptr(idx), ptr(p), ptr(w), stream_handle()), "cqs_draw_rows")
    return idx.cpu().numpy(), p.cpu().numpy()


def collapse_batch(states: torch.Tensor, n_bits: int, n_states: int, mask: int, values: np.ndarray,
                   scales: np.ndarray):
    _check_copies(states, n_bits, n_states)
    if np.shape(values) != (n_states,) or np.shape(scales) != (n_states,):
        raise ValueError("one value and one scale per copy")
    v = _dev_array(np.asarray(values, dtype=np.uint64), states.device)
    s = _dev_array(np.asarray(scales, dtype=np.float32), states.device)
    check(lib().cqs_collapse_batch(ptr(states), n_bits, n_states, int(mask), ptr(v), ptr(s), stream_handle()),
          "cqs_collapse_batch")


def device_mats(mats: np.ndarray, targets, device) -> tuple[torch.Tensor, int]:
    """A channel's 2^k x 2^k matrices (k = len(targets)) as a device complex64 buffer, and their count."""
    mats = np.asarray(mats, dtype=np.complex64)
    n_mats = int(mats.shape[0])
    d = 1 << len(targets)
    if mats.shape != (n_mats, d, d):
        raise ValueError(f"matrices of shape {mats.shape} for {len(targets)} target bits")
    return _dev_array(mats.view(np.float32), device), n_mats


def apply_select_batch_dev(states: torch.Tensor, n_bits: int, n_states: int, targets, mats_dev: torch.Tensor,
                           n_mats: int, select_dev: torch.Tensor, scales_dev: torch.Tensor):
    """Copy r <- scales[r] * mats[select[r]] on `targets` (matrix index bit m <-> flat bit
    targets[m]); matrices, selections (int32) and scales (float32) already on the device.  A copy
    whose selection is out of range is left unchanged."""
    _check_copies(states, n_bits, n_states)
    d = 1 << len(targets)
    if mats_dev.dtype != torch.float32 or mats_dev.numel() != n_mats * d * d * 2:
        raise ValueError("device matrices do not match n_mats and the targets")
    if select_dev.dtype != torch.int32 or select_dev.numel() < n_states:
        raise ValueError("one int32 selection per copy")
    if scales_dev.dtype != torch.float32 or scales_dev.numel() < n_states:
        raise ValueError("one float32 scale per copy")
    t = np.array(targets, dtype=np.int32)
    check(
        lib().cqs_apply_select_batch(ptr(states), n_bits, n_states, len(t), t.ctypes.data, ptr(mats_dev), n_mats,
                                     ptr(select_dev), ptr(scales_dev), stream_handle()),
        "cqs_apply_select_batch",
    )


def apply_select_batch(states: torch.Tensor, n_bits: int, n_states: int, targets, mats: np.ndarray,
                       select: np.ndarray, scales: np.ndarray):
    """apply_select_batch_dev with host arrays (uploaded here)."""
    m, n_mats = device_mats(mats, targets, states.device)
    sel = np.asarray(select, dtype=np.int32)
    if sel.shape != (n_states,) or np.shape(scales) != (n_states,):
        raise ValueError("one selection and one scale per copy")
    if sel.min() < 0 or sel.max() >= n_mats:
        raise ValueError("selection outside the channel's matrices")
    apply_select_batch_dev(states, n_bits, n_states, targets, m, n_mats, _dev_array(sel, states.device),
                           _dev_array(np.asarray(scales, dtype=np.float32), states.device))


def kraus_weights_batch_dev(states: torch.Tensor, n_bits: int, n_states: int, targets, mats_dev: torch.Tensor,
                            n_mats: int, ws: Workspace, to_host: bool = True):
    """||K_i psi_r||^2 in FP64 for every copy r and Kraus operator i ([n_states, n_mats]; a host
    array, or with to_host=False the device tensor, without a synchronisation)."""
    _check_copies(states, n_bits, n_states)
    d = 1 << len(targets)
    if mats_dev.dtype != torch.float32 or mats_dev.numel() != n_mats * d * d * 2:
        raise ValueError("device matrices do not match n_mats and the targets")
    t = np.array(targets, dtype=np.int32)
    out = torch.empty((n_states, n_mats), dtype=torch.float64, device=states.device)
    w = ws.get(lib().cqs_kraus_workspace_bytes(n_bits, n_states, len(t), n_mats))
    check(
        lib().cqs_kraus_weights_batch(ptr(states), n_bits, n_states, len(t), t.ctypes.data, ptr(mats_dev), n_mats,
                                      ptr(out), ptr(w), stream_handle()),
        "cqs_kraus_weights_batch",
    )
    return out.cpu().numpy() if to_host else out


def select_run_batch(states: torch.Tensor, n_bits: int, n_states: int, ops, mats_dev: torch.Tensor,
                     table_dev: torch.Tensor):
    """A run of per-copy mixture choices in one sweep of the copies.  ops: (k, targets, offset of
    the op's matrices in mats_dev in complex entries, their count, row of the op in table_dev)."""
    _check_copies(states, n_bits, n_states)
    if mats_dev.dtype != torch.float32 or table_dev.dtype != torch.int32 or table_dev.dim() != 2:
        raise ValueError("float32 matrix buffer and int32 [rows, copies] table expected")
    if table_dev.shape[1] != n_states or not table_dev.is_contiguous():
        raise ValueError("the table needs one contiguous column per copy")
    n = len(ops)
    k = np.array([o[0] for o in ops], dtype=np.int32)
    t = np.array([list(o[1]) + [0] * (2 - len(o[1])) for o in ops], dtype=np.int32).reshape(-1)
    mo = np.array([o[2] for o in ops], dtype=np.int32)
    nm = np.array([o[3] for o in ops], dtype=np.int32)
    row = np.array([o[4] for o in ops], dtype=np.int32)
    check(
        lib().cqs_select_run_batch(ptr(states), n_bits, n_states, n, k.ctypes.data, t.ctypes.data, mo.ctypes.data,
                                   nm.ctypes.data, row.ctypes.data, ptr(mats_dev), mats_dev.numel() // 2,
                                   ptr(table_dev), int(table_dev.shape[0]), stream_handle()),
        "cqs_select_run_batch",
    )


def select_then_kraus_weights(states: torch.Tensor, n_bits: int, n_states: int, bit_a: int, mats_a: torch.Tensor,
                              n_mats_a: int, sel_a: torch.Tensor, scl_a: torch.Tensor, bit_b: int,
                              mats_b: torch.Tensor, n_mats_b: int, ws: Workspace) -> torch.Tensor:
    """One sweep of the copies: channel A's chosen operators on bit_a, then channel B's weights on
    bit_b (device float64 [n_states, n_mats_b]).  Both channels act on one qubit."""
    _check_copies(states, n_bits, n_states)
    if mats_a.numel() != n_mats_a * 8 or mats_b.numel() != n_mats_b * 8:
        raise ValueError("one-qubit operator sets expected")
    if sel_a.dtype != torch.int32 or scl_a.dtype != torch.float32 or sel_a.numel() < n_states or scl_a.numel() < n_states:
        raise ValueError("one int32 selection and one float32 scale per copy")
    out = torch.empty((n_states, n_mats_b), dtype=torch.float64, device=states.device)
    w = ws.get(lib().cqs_kraus_workspace_bytes(n_bits, n_states, 1, n_mats_b))
    check(
        lib().cqs_select_then_kraus_weights(ptr(states), n_bits, n_states, int(bit_a), ptr(mats_a), n_mats_a,
                                            ptr(sel_a), ptr(scl_a), int(bit_b), ptr(mats_b), n_mats_b, ptr(out),
                                            ptr(w), stream_handle()),
        "cqs_select_then_kraus_weights",
    )
    return out


def kraus_choose_batch(weights_dev: torch.Tensor, uniforms_dev: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """Per copy, the Kraus operator the channel walk picks for its uniform and 1/sqrt(its weight),
    on the device (int32 selections, float32 scales)."""
    n_states, n_mats = weights_dev.shape
    if weights_dev.dtype != torch.float64 or not weights_dev.is_contiguous():
        raise ValueError("weights must be a contiguous float64 [copies, operators] tensor")
    if uniforms_dev.dtype != torch.float64 or not uniforms_dev.is_contiguous() or uniforms_dev.numel() < n_states:
        raise ValueError("one contiguous float64 uniform per copy")
    sel = torch.empty(n_states, dtype=torch.int32, device=weights_dev.device)
    scl = torch.empty(n_states, dtype=torch.float32, device=weights_dev.device)
    check(lib().cqs_kraus_choose_batch(ptr(weights_dev), ptr(uniforms_dev), n_states, n_mats, ptr(sel), ptr(scl),
                                       stream_handle()), "cqs_kraus_choose_batch")
    return sel, scl


def kraus_weights_batch(states: torch.Tensor, n_bits: int, n_states: int, targets, mats: np.ndarray,
                        ws: Workspace) -> np.ndarray:
    """kraus_weights_batch_dev with host matrices (uploaded here)."""
    m, n_mats = device_mats(mats, targets, states.device)
    return kraus_weights_batch_dev(states, n_bits, n_states, targets, m, n_mats, ws)


class Sampler:
    """Bit-exact NumPy-choice sampling of (marginal) probabilities of a state."""

    def __init__(self, device):
        self.device = device
        self.ws = Workspace(device)

    def marginal(self, state, n_qubits: int, is_dm: bool, axes) -> torch.Tensor:
        m = len(axes)
        out = torch.empty(1 << m, dtype=torch.float32, device=state.device)
        ax = np.array(axes, dtype=np.int32)
        check(
            lib().cqs_marginal_probs(ptr(state), n_qubits, int(is_dm), ax.ctypes.data, m, ptr(out), stream_handle()),
            "cqs_marginal_probs",
        )
        return out

    def draw(self, state, n_qubits: int, is_dm: bool, uniforms: np.ndarray, probs=None, axes=None) -> torch.Tensor:
        """Indices (device int64) of numpy's choice(2^m, p) for the given uniforms.

        probs: optional device float32 marginal (then n_qubits is its number of bits).
        axes: optional order of all qubits (a permutation) for sampling every qubit."""
        reps = int(uniforms.shape[0])
        w = self.ws.get(lib().cqs_sample_workspace_bytes(n_qubits))
        ax = None if axes is None else np.ascontiguousarray(axes, dtype=np.int32)
        if ax is not None and probs is None:
            if np.array_equal(ax, np.arange(n_qubits)):
                ax = None
            else:
                # transpose the probabilities once (contiguous reads and writes) instead of
                # gathering permuted amplitudes in every sampling kernel
                probs = self.permuted_probs(state, n_qubits, is_dm, ax)
                ax = None
        src = None if probs is not None else state
        check(
            lib().cqs_sample_prepare(ptr(src), ptr(probs), n_qubits, int(is_dm), ptr(ax), ptr(w), stream_handle()),
            "cqs_sample_prepare",
        )
        # staged through page-locked memory and queued: the host does not wait for the state's passes
        u = torch.from_numpy(np.ascontiguousarray(uniforms, dtype=np.float64)).pin_memory()
        u = u.to(self.device, non_blocking=True)
        idx = torch.empty(max(reps, 1), dtype=torch.int64, device=self.device)
        check(
            lib().cqs_sample_draw(
                ptr(src), ptr(probs), n_qubits, int(is_dm), ptr(ax), ptr(u), reps, ptr(idx), ptr(w), stream_handle()
            ),
            "cqs_sample_draw",
        )
        return idx[:reps]

    def permuted_probs(self, state, n_qubits: int, is_dm: bool, axes) -> torch.Tensor:
        """float32 probabilities of every basis state with the qubits in `axes` order."""
        if getattr(self, "_pbuf", None) is None or self._pbuf.numel() != (1 << n_qubits):
            self._pbuf = None
            self._pbuf = torch.empty(1 << n_qubits, dtype=torch.float32, device=self.device)
        ax = np.ascontiguousarray(axes, dtype=np.int32)
        check(lib().cqs_permuted_probs(ptr(state), n_qubits, int(is_dm), ax.ctypes.data, ptr(self._pbuf),
                                       stream_handle()), "cqs_permuted_probs")
        return self._pbuf

    def bits(self, idx: torch.Tensor, m: int) -> np.ndarray:
        reps = int(idx.shape[0])
        out = torch.empty(max(reps * m, 1), dtype=torch.uint8, device=self.device)
        check(lib().cqs_indices_to_bits(ptr(idx), reps, m, ptr(out), stream_handle()), "cqs_indices_to_bits")
        # one DMA copy into page-locked memory (torch's host allocator hands the block out again
        # once the returned array is dropped) instead of a staged copy into fresh pageable memory
        host = torch.empty(out.shape, dtype=torch.uint8, pin_memory=True)
        host.copy_(out, non_blocking=True)
        torch.cuda.current_stream().synchronize()
        return host[: reps * m].numpy().reshape(reps, m)
