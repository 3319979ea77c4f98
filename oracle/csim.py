"""Test infrastructure: ctypes front end of ``oracle/csim.c`` (complex128 kernels in C, OpenMP).

The same simulation as ``oracle.cirq_np.simulate_state_vector`` / ``simulate_density_matrix`` (the
matrices and Kraus lists come from that module's pinned tables), fast enough for the 28-qubit
grid circuit and the 12-qubit noisy density matrix whose results are committed as fixtures under
``tests/golden/``.  ``tests/test_oracle_csim.py`` checks it against the NumPy restatement.

Build: ``python -m oracle.csim`` (also run by ``__graft_entry__.build()``); the library goes to
``oracle/_build/libcsim.so``.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

from . import cirq_np as O

HERE = Path(__file__).resolve().parent
SRC = HERE / "csim.c"
LIB = HERE / "_build" / "libcsim.so"
_lib = None


def build() -> Path:
    LIB.parent.mkdir(exist_ok=True)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["gcc", "-O3", "-march=x86-64-v2", "-fopenmp", "-shared", "-fPIC", "-std=c11", str(SRC), "-o", str(tmp)]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists() or LIB.stat().st_mtime < SRC.stat().st_mtime:
            build()
        L = ctypes.CDLL(str(LIB))
        P = ctypes.c_void_p
        L.csim_apply.argtypes = [P, ctypes.c_int, ctypes.c_int, P, P]
        L.csim_apply_channel.argtypes = [P, ctypes.c_int, ctypes.c_int, P, ctypes.c_int, P]
        L.csim_expect_z.argtypes = [P, ctypes.c_int, ctypes.c_int, P, P]
        L.csim_expect_z.restype = None
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def apply_matrix(psi: np.ndarray, n: int, M: np.ndarray, axes) -> None:
    """psi (complex128[2^n], in place) <- M on qubit axes `axes` (big-endian)."""
    assert psi.dtype == np.complex128 and psi.flags.c_contiguous and psi.size == 1 << n
    M = np.ascontiguousarray(M, dtype=np.complex128)
    ax = np.ascontiguousarray(axes, dtype=np.int32)
    assert M.shape == (1 << len(ax), 1 << len(ax))
    rc = lib().csim_apply(_ptr(psi), n, len(ax), _ptr(ax), _ptr(M))
    if rc:
        raise ValueError(f"csim_apply failed ({rc}) for axes {list(axes)}")


def apply_channel(rho: np.ndarray, n: int, kraus, axes) -> None:
    """rho (complex128[4^n], in place) <- sum K rho K^dagger on qubit axes `axes`."""
    assert rho.dtype == np.complex128 and rho.flags.c_contiguous and rho.size == 1 << (2 * n)
    ks = np.ascontiguousarray(np.stack([np.asarray(k, dtype=np.complex128) for k in kraus]))
    ax = np.ascontiguousarray(axes, dtype=np.int32)
    rc = lib().csim_apply_channel(_ptr(rho), n, len(ax), _ptr(ax), len(ks), _ptr(ks))
    if rc:
        raise ValueError(f"csim_apply_channel failed ({rc}) for axes {list(axes)}")


def simulate_state_vector(ops, n: int, initial_state: int = 0) -> np.ndarray:
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[initial_state] = 1
    for kind, family, params, axes in ops:
        if kind != "gate":
            raise ValueError("state-vector runs take gates only")
        if family == "GlobalPhase":
            psi *= params["coefficient"]
            continue
        apply_matrix(psi, n, O.gate_unitary(family, params), axes)
    return psi


def simulate_density_matrix(ops, n: int, initial_state: int = 0) -> np.ndarray:
    rho = np.zeros(1 << (2 * n), dtype=np.complex128)
    rho[initial_state * (1 << n) + initial_state] = 1
    for kind, family, params, axes in ops:
        if kind == "gate":
            if family == "GlobalPhase":
                continue
            apply_channel(rho, n, [O.gate_unitary(family, params)], axes)
        else:
            apply_channel(rho, n, O.channel_kraus(family, params), axes)
    return rho


def expect_z(psi: np.ndarray, n_bits: int, zmasks) -> np.ndarray:
    """sum_i (-1)^popcount(i & z) |psi_i|^2 per flat-bit mask z (FP64)."""
    z = np.ascontiguousarray(zmasks, dtype=np.uint64)
    out = np.zeros(len(z), dtype=np.float64)
    lib().csim_expect_z(_ptr(psi), n_bits, len(z), _ptr(z), _ptr(out))
    return out


if __name__ == "__main__":
    print(build())
