"""Test infrastructure: a NumPy complex128 restatement of Cirq's simulation algorithm.

This package is the correctness oracle for the GPU engine.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import it, and only as the
checker / CPU baseline - never as part of the product path.  It follows the reference's NumPy
code function by function (citations in ``oracle/cirq_np.py``); gate matrices are restated from
the reference's eigen-component tables independently of ``cirq_amd``'s own gate code.
"""
