"""ctypes binding of ``libcirq_amd.so`` (the C ABI declared in include/cirq_amd.h).

The library is loaded from the package directory only; there is no fallback.  Errors from the
library are re-raised with Cirq's exception types: bad arguments -> ValueError, allocation
failure -> MemoryError, unsupported operations -> TypeError, HIP errors -> RuntimeError.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libcirq_amd.so"

CQS_OK, CQS_ERR_ARG, CQS_ERR_HIP, CQS_ERR_OOM, CQS_ERR_UNSUPPORTED = 0, 1, 2, 3, 4
OP_DENSE, OP_DIAG, OP_PAIRPHASE, OP_SCALE, OP_BLOCK = 1, 2, 3, 4, 5
OP_GRAD_DENSE, OP_GRAD_DIAG, OP_PERMUTE = 6, 7, 8
SEL_BOTH, SEL_PSI, SEL_PHI = 0, 1, 2
MAX_LOCAL = 16

# numpy mirror of cqs_op (16 x 32-bit words)
OP_DTYPE = np.dtype(
    [
        ("kind", np.int32),
        ("k", np.int32),
        ("mat", np.int32),
        ("slot", np.int32),
        ("t", np.int32, (6,)),
        ("cl_mask", np.int32),
        ("cl_val", np.int32),
        ("co_mask_lo", np.uint32),
        ("co_mask_hi", np.uint32),
        ("co_val_lo", np.uint32),
        ("co_val_hi", np.uint32),
    ]
)
assert OP_DTYPE.itemsize == 64

PASS_DTYPE = np.dtype(
    [
        ("L", np.int32),
        ("op_begin", np.int32),
        ("op_count", np.int32),
        ("max_k", np.int32),
        ("lpos", np.int32, (MAX_LOCAL,)),
        ("lpos_out", np.int32, (MAX_LOCAL,)),
    ]
)
assert PASS_DTYPE.itemsize == 144

BATCH_PASS_DTYPE = np.dtype(
    [
        ("n_bits", np.int32),
        ("L", np.int32),
        ("op_begin", np.int32),
        ("op_count", np.int32),
        ("max_k", np.int32),
        ("nouter", np.int32),
        ("pad", np.int32, (2,)),
        ("lpos", np.int32, (MAX_LOCAL,)),
        ("lpos_out", np.int32, (MAX_LOCAL,)),
        ("opos", np.int32, (48,)),
    ]
)
assert BATCH_PASS_DTYPE.itemsize == 352

BATCH_ITEM_DTYPE = np.dtype(
    [
        ("state", np.int64),
        ("state2", np.int64),
        ("mats", np.int64),
        ("partial", np.int64),
        ("pass", np.int32),
        ("tile_begin", np.int32),
    ]
)
assert BATCH_ITEM_DTYPE.itemsize == 40

BATCH_TERMS_DTYPE = np.dtype(
    [
        ("state", np.int64),
        ("state2", np.int64),
        ("out", np.int64),
        ("n_bits", np.int32),
        ("term_begin", np.int32),
        ("n_terms", np.int32),
        ("block_begin", np.int32),
    ]
)
assert BATCH_TERMS_DTYPE.itemsize == 40

_lib = None

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_sz = ctypes.c_size_t
_f32 = ctypes.c_float

_SIGNATURES = {
    "cqs_abi_version": ([], ctypes.c_int),
    "cqs_last_error": ([], ctypes.c_char_p),
    "cqs_device_info": ([_vp, _vp], ctypes.c_int),
    "cqs_init_basis": ([_vp, _i32, _u64, _vp], ctypes.c_int),
    "cqs_run_passes": ([_vp, _i32, _vp, _i32, _vp, _vp, _vp, _vp], ctypes.c_int),
    "cqs_run_passes_from_basis": ([_vp, _i32, _u64, _vp, _i32, _vp, _vp, _vp, _vp], ctypes.c_int),
    "cqs_program_validate": ([_i32, _vp, _i32, _vp, _i32, _i64, _vp, _i64, _i32], ctypes.c_int),
    "cqs_expect_workspace_bytes": ([_i32], _sz),
    "cqs_expect_paulis": ([_vp, _i32, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "cqs_expect_paulis_noalloc": ([_vp, _i32, _i32, _i32, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "cqs_apply_paulisum": ([_vp, _vp, _i32, _i32, _vp, _vp, _vp, _vp], ctypes.c_int),
    "cqs_norm2": ([_vp, _i32, _vp, _vp, _vp], ctypes.c_int),
    "cqs_inner": ([_vp, _vp, _i32, _vp, _vp, _vp], ctypes.c_int),
    "cqs_run_passes_specialized": ([_vp, _i32, ctypes.c_int64, _vp, _i32, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "cqs_adjoint_passes_specialized": ([_vp, _vp, _i32, _vp, _i32, _vp, _vp, _vp, _vp, ctypes.c_int64, _vp, _vp],
                                       ctypes.c_int),
    "cqs_scale": ([_vp, _i32, _f32, _f32, _vp], ctypes.c_int),
    "cqs_collapse": ([_vp, _i32, _u64, _u64, _f32, _f32, _vp], ctypes.c_int),
    "cqs_sample_workspace_bytes": ([_i32], _sz),
    "cqs_sample_prepare": ([_vp, _vp, _i32, _i32, _vp, _vp, _vp], ctypes.c_int),
    "cqs_sample_draw": ([_vp, _vp, _i32, _i32, _vp, _vp, _i64, _vp, _vp, _vp], ctypes.c_int),
    "cqs_marginal_probs": ([_vp, _i32, _i32, _vp, _i32, _vp, _vp], ctypes.c_int),
    "cqs_marginal_probs_batch": ([_vp, _i32, _i32, _i32, _vp, _i32, _vp, _vp], ctypes.c_int),
    "cqs_draw_rows_workspace_bytes": ([_i32, _i32], _sz),
    "cqs_draw_rows": ([_vp, _i32, _i32, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "cqs_collapse_batch": ([_vp, _i32, _i32, _u64, _vp, _vp, _vp], ctypes.c_int),
    "cqs_apply_select_batch": ([_vp, _i32, _i32, _i32, _vp, _vp, _i32, _vp, _vp, _vp], ctypes.c_int),
    "cqs_kraus_workspace_bytes": ([_i32, _i32, _i32, _i32], _sz),
    "cqs_kraus_weights_batch": ([_vp, _i32, _i32, _i32, _vp, _vp, _i32, _vp, _vp, _vp], ctypes.c_int),
    "cqs_kraus_choose_batch": ([_vp, _vp, _i32, _i32, _vp, _vp, _vp], ctypes.c_int),
    "cqs_select_then_kraus_weights": ([_vp, _i32, _i32, _i32, _vp, _i32, _vp, _vp, _i32, _vp, _i32, _vp, _vp, _vp],
                                      ctypes.c_int),
    "cqs_select_run_batch": ([_vp, _i32, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _i32, _vp],
                             ctypes.c_int),
    "cqs_indices_to_bits": ([_vp, _i64, _i32, _vp, _vp], ctypes.c_int),
    "cqs_permuted_probs": ([_vp, _i32, _i32, _vp, _vp, _vp], ctypes.c_int),
    "cqs_adjoint_passes": ([_vp, _vp, _i32, _vp, _i32, _vp, _vp, _vp, _vp, _i64, _vp], ctypes.c_int),
    "cqs_grad_reduce": ([_vp, _i32, _i64, _vp, _vp], ctypes.c_int),
    "cqs_batch_validate": ([_vp, _i32, _i32, _vp, _i32, _i32, _i64, _i64, _i32], ctypes.c_int),
    "cqs_batch_init": ([_vp, _vp, _i32, _i64, _vp, _i32, _vp], ctypes.c_int),
    "cqs_batch_passes": ([_vp, _vp, _i32, _i64, _i32, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "cqs_batch_adjoint": ([_vp, _vp, _i32, _i64, _i32, _vp, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "cqs_batch_apply_paulisum": ([_vp, _vp, _i32, _i64, _vp, _vp, _vp, _i32, _vp], ctypes.c_int),
    "cqs_batch_expect": ([_vp, _vp, _i32, _i64, _vp, _vp, _vp, _i32, _vp], ctypes.c_int),
    "cqs_segment_sum": ([_vp, _vp, _i32, _i32, _vp, _vp], ctypes.c_int),
}


def exported_symbols() -> list[str]:
    return list(_SIGNATURES)


def lib():
    """Loads the library (once).  Raises RuntimeError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `python cirq_amd/_build.py`"
        )
    l = ctypes.CDLL(str(LIB_PATH))
    for name, (args, res) in _SIGNATURES.items():
        f = getattr(l, name)
        f.argtypes = args
        f.restype = res
    _lib = l
    return l


def check(rc: int, what: str = ""):
    if rc == CQS_OK:
        return
    msg = lib().cqs_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == CQS_ERR_ARG:
        raise ValueError(text)
    if rc == CQS_ERR_OOM:
        raise MemoryError(text)
    if rc == CQS_ERR_UNSUPPORTED:
        raise TypeError(text)
    raise RuntimeError(text)


def ptr(x) -> int | None:
    """Raw pointer of a torch tensor / numpy array / int / None."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if isinstance(x, np.ndarray):
        return x.ctypes.data
    return x.data_ptr()
