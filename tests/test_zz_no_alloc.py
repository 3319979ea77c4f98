"""cqs_expect_paulis_noalloc (static device scratch, no allocation) against cqs_expect_paulis
(caller workspace), through the C ABI: bit-identical values on two calls in a row, the second on
a smaller input, and both against an FP64 NumPy sum."""

import numpy as np
import pytest
import torch

from cirq_amd import _lib
from cirq_amd._lib import check, lib, ptr

pytestmark = pytest.mark.gpu


def _terms(rng, n, count):
    xs, zs, cs = [], [], []
    for t in range(count):
        x = 0 if t % 3 == 0 else int(rng.randint(0, 2**n))  # a third Z-only (the k_expect_z path)
        z = int(rng.randint(0, 2**n))
        xs.append(x)
        zs.append(z)
        cs.append(complex(rng.randn(), rng.randn()))
    return (np.array(xs, dtype=np.uint64), np.array(zs, dtype=np.uint64),
            np.array([[c.real, c.imag] for c in cs], dtype=np.float64).reshape(-1))


def _both(st, n, is_dm, xm, zm, cf):
    T = len(xm)
    dev = st.device
    ws = torch.empty(int(lib().cqs_expect_workspace_bytes(T)), dtype=torch.uint8, device=dev)
    a = torch.zeros(2 * T, dtype=torch.float64, device=dev)
    b = torch.full((2 * T,), np.nan, dtype=torch.float64, device=dev)
    s = torch.cuda.current_stream().cuda_stream
    check(lib().cqs_expect_paulis(ptr(st), n, is_dm, T, xm.ctypes.data, zm.ctypes.data, cf.ctypes.data, ptr(a),
                                  ptr(ws), s), "cqs_expect_paulis")
    check(lib().cqs_expect_paulis_noalloc(ptr(st), n, is_dm, T, xm.ctypes.data, zm.ctypes.data, cf.ctypes.data,
                                          ptr(b), s), "cqs_expect_paulis_noalloc")
    torch.cuda.synchronize()
    return a.cpu().numpy().view(np.complex128), b.cpu().numpy().view(np.complex128)


def _numpy_sv(psi, xm, zm, cf):
    idx = np.arange(psi.size, dtype=np.uint64)
    p = psi.astype(np.complex128)
    out = []
    for x, z, k in zip(xm, zm, range(len(xm))):
        c = complex(cf[2 * k], cf[2 * k + 1])
        sign = 1 - 2 * (np.array([bin(int(v)).count("1") for v in (idx & z)]) & 1)
        out.append(c * np.sum(np.conj(p[(idx ^ x).astype(np.int64)]) * sign * p))
    return np.array(out)


def test_noalloc_expectations_match_workspace_entry_point(gpu):
    _lib.lib()
    rng = np.random.RandomState(7)
    for n, count in ((20, 60), (12, 7)):  # the second call is smaller
        psi = (rng.randn(2**n) + 1j * rng.randn(2**n)).astype(np.complex64)
        psi /= np.linalg.norm(psi)
        st = torch.from_numpy(psi).to(gpu)
        xm, zm, cf = _terms(rng, n, count)
        a, b = _both(st, n, 0, xm, zm, cf)
        assert np.array_equal(a, b), f"n={n}: {np.max(np.abs(a - b))}"
        if n <= 12:
            np.testing.assert_allclose(b, _numpy_sv(psi, xm, zm, cf), atol=1e-12)
    # a density matrix on 5 qubits (row bits above column bits)
    n = 5
    v = rng.randn(2**n, 3) + 1j * rng.randn(2**n, 3)
    rho = (v @ v.conj().T)
    rho = (rho / np.trace(rho)).astype(np.complex64)
    st = torch.from_numpy(rho.reshape(-1)).to(gpu)
    xm, zm, cf = _terms(rng, n, 9)
    a, b = _both(st, n, 1, xm, zm, cf)
    assert np.array_equal(a, b)


def test_noalloc_rejects_bad_arguments(gpu):
    _lib.lib()
    st = torch.zeros(16, dtype=torch.complex64, device=gpu)
    out = torch.zeros(2, dtype=torch.float64, device=gpu)
    xm = np.array([1 << 10], dtype=np.uint64)  # beyond 4 qubits
    zm = np.zeros(1, dtype=np.uint64)
    cf = np.array([1.0, 0.0])
    rc = lib().cqs_expect_paulis_noalloc(ptr(st), 4, 0, 1, xm.ctypes.data, zm.ctypes.data, cf.ctypes.data, ptr(out),
                                         torch.cuda.current_stream().cuda_stream)
    assert rc == _lib.CQS_ERR_ARG
    assert b"beyond" in lib().cqs_last_error()
