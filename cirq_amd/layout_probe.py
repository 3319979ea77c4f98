"""Streaming-rate probe of the pass kernel's tile layouts (a diagnostic, not on the simulation path).

`python -m cirq_amd.layout_probe [rows cols depth]` prints one JSON object: the device-to-device
copy rate of a rows x cols-qubit state, the rate of load/store-only passes (k_pass with one SCALE
by 1) for 12-bit tiles whose 8 upper bits sit in different ranges of the flat index, and, for the
random grid circuit compiled with 12- and 13-bit tiles, the sum of such passes on the program's
own pass layouts (its streaming floor) beside the program's real pass time.
"""

from __future__ import annotations

import json
import sys

import numpy as np


def layout_probe(rows: int = 5, cols: int = 6, depth: int = 40) -> dict:
    """How fast HBM streams for the access patterns the pass kernel makes, on the rows x cols state.

    Times (HIP events) a device-to-device copy of the state and passes that only load and store it
    (one SCALE by 1) with a 12-bit tile: 4 contiguous low bits plus 8 bits placed in different
    ranges of the flat index.  Bits >= 18 are 2 MiB apart or more in memory.  Then, for the grid
    circuit compiled with 12- and 13-bit tiles, the sum of such load/store-only passes on the
    program's own pass layouts (its streaming floor) beside the program's real pass time."""
    import torch

    from cirq_amd import engine, generators, tape

    n = rows * cols
    dev = engine.require_gpu()
    st = engine.new_state(n, dev)
    gb = 2 * 8 * 2**n / 1e9

    def timed(fn, reps=4):
        fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    def load_store(L, lpos):
        one = tape.LOp("scale", (), np.complex128(1.0))
        dp = engine.DeviceProgram(tape.PackedProgram([tape.Pass(L, tuple(lpos), [one])], n), dev)
        return timed(lambda: dp.run(st))

    out = {"n_qubits": n, "state_gb": gb / 2}
    y = torch.empty_like(st)
    ms = timed(lambda: y.copy_(st))
    out["device_copy"] = {"ms": ms, "tb_s": gb / ms}
    del y
    torch.cuda.empty_cache()
    low = list(range(4))
    pick = lambda lo, hi, seed: sorted(np.random.RandomState(seed).choice(np.arange(lo, hi), 8, replace=False).tolist())
    layouts = {
        "contiguous 0-11": [list(range(12))],
        "4 low + 8 in [4,18)": [low + pick(4, 18, s) for s in range(3)],
        "4 low + 8 in [18,n)": [low + pick(18, n, s) for s in range(3)],
        "4 low + top 8": [low + list(range(n - 8, n))],
    }
    out["load_store_L12"] = {}
    for name, lps in layouts.items():
        ts = [load_store(12, lp) for lp in lps]
        out["load_store_L12"][name] = {"ms": float(np.mean(ts)), "tb_s": gb / float(np.mean(ts)), "layouts": lps}
    circuit = generators.grid_circuit(rows, cols, depth, seed=0)
    lw = tape.Lowering(sorted(circuit.all_qubits()), dm=False)
    lw.lower_circuit(circuit)
    out["grid_program"] = {}
    for L in (12, 13):
        p = tape.compile_ops(lw.ops, n, L=L, coalesce=4, kmax=4)
        dp = engine.DeviceProgram(p, dev)
        real = timed(lambda: dp.run(st), reps=2)
        floor = sum(load_store(int(r["L"]), [int(x) for x in r["lpos"][: int(r["L"])]]) for r in p.pass_arr)
        high = [sum(1 for x in r["lpos"][: int(r["L"])] if int(x) >= 18) for r in p.pass_arr]
        out["grid_program"][f"L{L}"] = {"passes": int(p.n_passes), "pass_ms": real, "load_store_floor_ms": floor,
                                       "mean_local_bits_ge_18": float(np.mean(high))}
    return out


if __name__ == "__main__":
    a = [int(x) for x in sys.argv[1:4]]
    print(json.dumps(layout_probe(*a)))
