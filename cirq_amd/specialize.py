"""Pass programs compiled to straight-line kernels.

The pass kernel (`k_pass`, `csrc/tile.hpp`) interprets a pass's op list: a loop that dispatches on
each op's kind and bit positions while the tile sits in registers.  The dispatch costs more than
the branches: an op's results are formed while its operands are live, so they land in fresh
registers, and every trip of the loop copies all of a thread's amplitudes back into the registers
the loop keeps them in (64 moves per op at 16 amplitudes per thread, measured in the ISA of
`k_pass<13>`).  Here each pass becomes its own kernel that calls the same device functions in the
pass's order, with the kind, the register positions, the matrix offsets and the outer-bit tests
as constants; the compiler then renames instead of copying and resolves every dispatch, down to
the MFMA block width and whether its column blocks are whole.

`specialize(packed)` writes the kernels' HIP source, compiles it with hipcc for gfx950 into a
shared object under the cache directory (reused when the same source was compiled before) and
returns the host launch handles, which `cqs_run_passes_specialized` launches with the same
checked pass arguments as the interpreter.  Compiling takes seconds to a minute (parallel
hipcc, a few kernels per file), so it pays for circuits that run many times (sampling loops,
parameter sweeps, benchmarks) on states large enough for the passes to be compute-bound.

Cirq has no counterpart: its NumPy path has no kernels to compile.
"""

from __future__ import annotations

import concurrent.futures
import ctypes
import hashlib
import os
import subprocess
from pathlib import Path

import numpy as np

from . import _build, _lib

CSRC = Path(__file__).resolve().parent / "csrc"
CACHE_DIR = Path(os.environ.get("CIRQ_AMD_SPECIALIZE_CACHE", Path(__file__).resolve().parent / "_specialized"))
KERNELS_PER_FILE = 4

_OP_DENSE, _OP_DIAG, _OP_PAIR, _OP_SCALE, _OP_BLOCK, _OP_PERMUTE = (
    _lib.OP_DENSE, _lib.OP_DIAG, _lib.OP_PAIRPHASE, _lib.OP_SCALE, _lib.OP_BLOCK, _lib.OP_PERMUTE)


def _shape(L: int) -> tuple[int, int]:
    """Threads and amplitudes per thread of a tile of 2^L amplitudes (tile.hpp TileShape)."""
    nt = 256 if L <= 12 else (1 << L) // 16
    per = 1 if L <= 8 else (1 << L) // nt
    return nt, per


def _is_register_dense(op, nt: int, per: int) -> bool:
    """tile.hpp is_register_dense: k <= 3 targets, all register bits, ascending."""
    tb, nr = nt.bit_length() - 1, per.bit_length() - 1
    k = int(op["k"])
    if (int(op["kind"]) & 0xFF) != _OP_DENSE or k < 1 or k > 3 or k > nr:
        return False
    t = [int(x) for x in op["t"][:k]]
    return t[0] >= tb and all(t[m] > t[m - 1] for m in range(1, k))


def _mixing(op, nt: int, per: int) -> bool:
    kind = int(op["kind"]) & 0xFF
    return (kind == _OP_DENSE and not _is_register_dense(op, nt, per)) or kind == _OP_BLOCK


def _outer_guard(op) -> str:
    """`if (...) ` testing the op's outer controls against the tile's outer bits, or ''."""
    mask = int(op["co_mask_lo"]) | (int(op["co_mask_hi"]) << 32)
    if not mask:
        return ""
    val = int(op["co_val_lo"]) | (int(op["co_val_hi"]) << 32)
    return f"if ((base & 0x{mask:x}ull) == 0x{val:x}ull) "


def _variant_expr(op) -> str:
    """The op's matrix variant from the tile's outer bits (tile.hpp variant_of), as constants."""
    code = int(op["t"][5])
    cnt = (code >> 24) & 3
    if not cnt:
        return ""
    parts = [f"((int)((base >> {(code >> (7 * i)) & 127}) & 1ull) << {i})" for i in range(cnt)]
    return "(" + " | ".join(parts) + ")"


def _op_constant(name: str, op) -> str:
    """The op record as a namespace-scope constant, so that the device functions fold its fields
    (targets, controls, matrix offset) instead of loading them."""
    t = ", ".join(str(int(x)) for x in op["t"])
    return (f"static constexpr cqs_op {name} = {{{int(op['kind'])}, {int(op['k'])}, {int(op['mat'])}, {int(op['slot'])}, "
            f"{{{t}}}, {int(op['cl_mask'])}, {int(op['cl_val'])}, {int(op['co_mask_lo'])}u, {int(op['co_mask_hi'])}u, "
            f"{int(op['co_val_lo'])}u, {int(op['co_val_hi'])}u}};\n")


def _mixing_call(op, ref: str, L: int, tile: str, scaled: bool = False) -> str:
    """tile.hpp run_mixing_op for one op with its kind, width and matrix offset resolved: an MFMA
    block or a VALU dense gate on the LDS tile `tile`.  scaled: the tile already holds amplitudes
    times P.mscale, so an MFMA block splits and stores them as they are."""
    kind, k, mat = int(op["kind"]) & 0xFF, int(op["k"]), int(op["mat"])
    var = _variant_expr(op)
    if kind == _OP_BLOCK:
        a = f"P.mats + {mat}" + (f" + {var} * {2 << (2 * k)}" if var else "")
        full = "true" if L - k >= 5 else "false"
        scale = "1.f, 1.f" if scaled else "P.mscale, 1.f / P.mscale"
        return (f"tile_block_mfma<{k}, NT, {full}>({tile}, L, {ref}, reinterpret_cast<const half8_t*>({a}), "
                f"{scale});")
    m = f"scalar_ptr(P.mats) + {mat}" + (f" + {var} * {1 << (2 * k)}" if var else "")
    if k in (1, 2):
        return f"tile_dense<{k}, NT>({tile}, L, {ref}, {m});"
    if k == 3:
        return f"stage_matrix<NT>(mstage, {m}, 64); __syncthreads(); tile_dense<3, NT>({tile}, L, {ref}, mstage);"
    raise ValueError(f"mixing op of kind {kind} on {k} bits has no specialised form")


def pass_source(index: int, ps, ops: np.ndarray) -> str:
    """HIP source of one pass as `extern "C" __global__ void k_spec_<index>(float2*, TileProg)`,
    the same steps as tile.hpp pass_tile / run_program for this pass's ops."""
    L = int(ps["L"])
    nt, per = _shape(L)
    tb = nt.bit_length() - 1
    first, count = int(ps["op_begin"]), int(ps["op_count"])
    consts = "".join(_op_constant(f"kS{index}_{j}", ops[first + j]) for j in range(count))
    # Every op of a pass is linear, and scaling by the power of two P.mscale is exact, so a pass
    # with MFMA blocks scales its amplitudes once after the load and once before the store instead
    # of around every block (same results: FP32 rounding does not depend on a power-of-two scale).
    scaled = any((int(ops[first + j]["kind"]) & 0xFF) == _OP_BLOCK for j in range(count))
    body = []
    j = 0
    while j < count:
        op = ops[first + j]
        kind = int(op["kind"]) & 0xFF
        if not _mixing(op, nt, per):
            ref = f"kS{index}_{j}"
            guard = _outer_guard(op)
            if kind == _OP_PERMUTE:
                body.append(f"  __syncthreads();\n  regs_to_tile_permuted<PER, NT>(v, tile, L, scalar_ptr(P.aux) + {int(op['t'][0])}, act);\n"
                            "  __syncthreads();\n  tile_to_regs<PER, NT>(v, tile, act);")
            elif kind == _OP_DENSE:
                k = int(op["k"])
                r = [int(x) - tb for x in op["t"][:k]] + [0, 0]
                var = _variant_expr(op)
                mat = f"scalar_ptr(P.mats) + {int(op['mat'])}" + (f" + {var} * {1 << (2 * k)}" if var else "")
                ctl = "true" if int(op["cl_mask"]) else "false"
                body.append(f"  {guard}reg_dense_k<PER, NT, {k}, {r[0]}, {r[1]}, {r[2]}, {ctl}>(v, {mat}, "
                            f"{int(op['cl_mask'])}, {int(op['cl_val'])});")
            elif kind == _OP_DIAG:
                body.append(f"  {guard}ew_diag<PER, NT>(v, P, {ref}, base);")
            elif kind == _OP_PAIR:
                body.append(f"  {guard}ew_pairs<PER, NT>(v, P, {ref}, base);")
            elif kind == _OP_SCALE:
                body.append(f"  {guard}ew_scale<PER>(v, P, {ref});")
            else:
                raise ValueError(f"pass {index}: op kind {kind} has no specialised form")
            j += 1
            continue
        body.append("  regs_to_tile<PER, NT>(v, tile, act);")
        while j < count and _mixing(ops[first + j], nt, per):
            op = ops[first + j]
            guard = _outer_guard(op)
            call = _mixing_call(op, f"kS{index}_{j}", L, "tile", scaled)
            body.append(f"  __syncthreads();\n  {guard}{{ {call} }}" if guard else f"  __syncthreads();\n  {call}")
            j += 1
        body.append("  __syncthreads();\n  tile_to_regs<PER, NT>(v, tile, act);")
    lines = "\n".join(body)
    one = "P.mscale" if scaled else "1.f"
    load = "v[e] = psi[base + off_lo + off_hi[e]];"
    if scaled:
        load = ("{ const float2 a = psi[base + off_lo + off_hi[e]]; "
                "v[e] = make_float2(a.x * P.mscale, a.y * P.mscale); }")
    store = "make_float2(v[e].x * inv, v[e].y * inv)" if scaled else "v[e]"
    inv = "\n    const float inv = 1.f / P.mscale;" if scaled else ""
    return consts + f"""
extern "C" __global__ __launch_bounds__({nt}) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_spec_{index}(
    float2* __restrict__ psi, TileProg P) {{
  constexpr int L = {L}, NT = {nt}, PER = {per};
  extern __shared__ float2 lds[];
  uint32_t b = blockIdx.x;  // XCD-aware tile order, as k_pass
  const uint32_t G = gridDim.x;
  if ((G & 7) == 0) b = (b & 7) * (G >> 3) + (b >> 3);
  const uint64_t base = deposit_bounded<31>((uint64_t)b, P.opos, P.nouter);
  float2* tile = lds;
  float2* mstage = lds + tile_words(L);
  sop_t* ops = scalar_ptr(P.ops);
  const bool act = L >= 8 || (int)threadIdx.x < (1 << L);
  (void)ops;
  (void)mstage;
  if ((base & P.wmask) != P.wval) return;  // nobody reads this tile before the next pass stores it
  if ((base & P.zmask) != P.zval) {{  // a tile of zeros stays zero (ops map zero to zero)
    store_zero_tile<L, PER, NT>(psi, P.lpos_out, base, act);
    return;
  }}
  warm_program<NT>(P);
  float2 v[PER];
  {{
    uint64_t off_lo, off_hi[PER];
    tile_offsets<L>(P.lpos, off_lo, off_hi);
    if (act) {{
      if (P.basis >= 0) {{
#pragma unroll
        for (int e = 0; e < PER; ++e)
          v[e] = make_float2(base + off_lo + off_hi[e] == (uint64_t)P.basis ? {one} : 0.f, 0.f);
      }} else {{
#pragma unroll
        for (int e = 0; e < PER; ++e) {load}
      }}
    }}
  }}
{lines}
  {{
    uint64_t off_lo, off_hi[PER];
    tile_offsets<L>(P.lpos_out, off_lo, off_hi);{inv}
    if (act) {{
#pragma unroll
      for (int e = 0; e < PER; ++e) psi[base + off_lo + off_hi[e]] = {store};
    }}
  }}
}}
"""


def _adj_register_dense(op, nt: int, per: int) -> bool:
    """adjoint.hpp adj_register_dense: DENSE or GRAD_DENSE, k <= 3 register bits, ascending."""
    tb, nr = nt.bit_length() - 1, per.bit_length() - 1
    kind, k = int(op["kind"]) & 0xFF, int(op["k"])
    if kind not in (_OP_DENSE, _lib.OP_GRAD_DENSE) or k < 1 or k > 3 or k > nr:
        return False
    t = [int(x) for x in op["t"][:k]]
    return t[0] >= tb and all(t[m] > t[m - 1] for m in range(1, k))


def _adj_mixing(op, nt: int, per: int) -> bool:
    kind = int(op["kind"]) & 0xFF
    if kind == _OP_BLOCK:
        return True
    return kind in (_OP_DENSE, _lib.OP_GRAD_DENSE) and not _adj_register_dense(op, nt, per)


def _guard_cond(op) -> str:
    """The op's outer-control test as a C condition ('true' without outer controls)."""
    mask = int(op["co_mask_lo"]) | (int(op["co_mask_hi"]) << 32)
    if not mask:
        return "true"
    val = int(op["co_val_lo"]) | (int(op["co_val_hi"]) << 32)
    return f"(base & 0x{mask:x}ull) == 0x{val:x}ull"


def adjoint_pass_source(index: int, ps, ops: np.ndarray) -> str:
    """HIP source of one reverse-sweep pass as `extern "C" __global__ void k_aspec_<index>(float2*
    psi, float2* phi, TileProg, double* partial, int64_t stride)`, the same steps as adjoint.hpp
    adjoint_tile for this pass's ops."""
    L = int(ps["L"])
    nt, per = _shape(L)
    tb = nt.bit_length() - 1
    first, count = int(ps["op_begin"]), int(ps["op_count"])
    consts = "".join(_op_constant(f"kA{index}_{j}", ops[first + j]) for j in range(count))
    body = []
    j = 0
    sel_states = {0: ("va", "vb"), 1: ("va",), 2: ("vb",)}
    while j < count:
        op = ops[first + j]
        kind, sel = int(op["kind"]) & 0xFF, (int(op["kind"]) >> 8) & 3
        ref = f"kA{index}_{j}"
        if not _adj_mixing(op, nt, per):
            cond = _guard_cond(op)
            if kind in (_lib.OP_GRAD_DIAG, _lib.OP_GRAD_DENSE):
                if kind == _lib.OP_GRAD_DIAG:
                    call = f"grad_reg_diag<PER, NT>(va, vb, P, {ref}, base)"
                else:
                    k = int(op["k"])
                    r = [int(x) - tb for x in op["t"][:k]] + [0, 0]
                    call = f"grad_reg_dense_k<PER, {k}, {r[0]}, {r[1]}, {r[2]}>(va, vb, scalar_ptr(P.mats) + {int(op['mat'])})"
                body.append(f"  {{\n    double g = 0.0;\n    if (({cond}) && act) g = {call};\n"
                            f"    write_wave_partial<NT>(g, partial, stride, tile, {int(op['slot'])});\n  }}")
                j += 1
                continue
            guard = "" if cond == "true" else f"if ({cond}) "
            states = sel_states[sel]
            for st in states:
                if kind == _OP_DENSE:
                    k = int(op["k"])
                    r = [int(x) - tb for x in op["t"][:k]] + [0, 0]
                    var = _variant_expr(op)
                    mat = f"scalar_ptr(P.mats) + {int(op['mat'])}" + (f" + {var} * {1 << (2 * k)}" if var else "")
                    ctl = "true" if int(op["cl_mask"]) else "false"
                    body.append(f"  {guard}reg_dense_k<PER, NT, {k}, {r[0]}, {r[1]}, {r[2]}, {ctl}>({st}, {mat}, "
                                f"{int(op['cl_mask'])}, {int(op['cl_val'])});")
                elif kind == _OP_DIAG:
                    body.append(f"  {guard}ew_diag<PER, NT>({st}, P, {ref}, base);")
                elif kind == _OP_PAIR:
                    body.append(f"  {guard}ew_pairs<PER, NT>({st}, P, {ref}, base);")
                elif kind == _OP_SCALE:
                    body.append(f"  {guard}ew_scale<PER>({st}, P, {ref});")
                else:
                    raise ValueError(f"reverse pass {index}: op kind {kind} has no specialised form")
            j += 1
            continue
        body.append("  regs_to_tile<PER, NT>(va, tA, act);\n  regs_to_tile<PER, NT>(vb, tB, act);")
        while j < count and _adj_mixing(ops[first + j], nt, per):
            op = ops[first + j]
            kind, sel = int(op["kind"]) & 0xFF, (int(op["kind"]) >> 8) & 3
            ref = f"kA{index}_{j}"
            cond = _guard_cond(op)
            body.append("  __syncthreads();")
            if kind == _lib.OP_GRAD_DENSE:
                k = int(op["k"])
                body.append(f"  {{\n    double g = 0.0;\n    if ({cond}) g = grad_dense<{k}, NT>(tA, tB, L, {ref}, "
                            f"scalar_ptr(P.mats) + {int(op['mat'])});\n"
                            f"    write_wave_partial<NT>(g, partial, stride, tile, {int(op['slot'])});\n  }}")
            else:
                inner = []
                if sel != 2:
                    inner.append(_mixing_call(op, ref, L, "tA"))
                if sel == 0:
                    inner.append("__syncthreads();")
                if sel != 1:
                    inner.append(_mixing_call(op, ref, L, "tB"))
                text = " ".join(inner)
                body.append(f"  {text}" if cond == "true" else f"  if ({cond}) {{ {text} }}")
            j += 1
        body.append("  __syncthreads();\n  tile_to_regs<PER, NT>(va, tA, act);\n  tile_to_regs<PER, NT>(vb, tB, act);")
    lines = "\n".join(body)
    return consts + f"""
extern "C" __global__ __launch_bounds__({nt}) __attribute__((amdgpu_waves_per_eu(2, 8))) void k_aspec_{index}(
    float2* __restrict__ psi, float2* __restrict__ phi, TileProg P, double* __restrict__ partial, int64_t stride) {{
  constexpr int L = {L}, NT = {nt}, PER = {per};
  extern __shared__ float2 lds[];
  const int64_t tile = blockIdx.x;
  const uint64_t base = deposit_bounded<31>((uint64_t)blockIdx.x, P.opos, P.nouter);
  float2* tA = lds;
  float2* tB = lds + tile_words(L);
  float2* mstage = lds + 2 * tile_words(L);
  sop_t* ops = scalar_ptr(P.ops);
  const bool act = L >= 8 || (int)threadIdx.x < (1 << L);
  (void)ops;
  (void)mstage;
  warm_program<NT>(P);
  float2 va[PER], vb[PER];
  {{
    uint64_t off_lo, off_hi[PER];
    tile_offsets<L>(P.lpos, off_lo, off_hi);
    if (act) {{
#pragma unroll
      for (int e = 0; e < PER; ++e) {{
        va[e] = psi[base + off_lo + off_hi[e]];
        vb[e] = phi[base + off_lo + off_hi[e]];
      }}
    }} else {{
#pragma unroll
      for (int e = 0; e < PER; ++e) va[e] = vb[e] = make_float2(0.f, 0.f);
    }}
  }}
{lines}
  {{
    uint64_t off_lo, off_hi[PER];
    tile_offsets<L>(P.lpos, off_lo, off_hi);
    if (act) {{
#pragma unroll
      for (int e = 0; e < PER; ++e) {{
        psi[base + off_lo + off_hi[e]] = va[e];
        phi[base + off_lo + off_hi[e]] = vb[e];
      }}
    }}
  }}
}}
"""


def program_sources(packed, adjoint: bool = False) -> list[str]:
    """The program's kernels (reverse-sweep kernels for an adjoint program), KERNELS_PER_FILE to a
    translation unit."""
    head = f'#include "{CSRC / ("adjoint.hpp" if adjoint else "tile.hpp")}"\nusing namespace cqs;\n'
    ops = packed.ops_arr
    gen = adjoint_pass_source if adjoint else pass_source
    kernels = [gen(i, ps, ops) for i, ps in enumerate(packed.pass_arr)]
    per = 1 if adjoint else KERNELS_PER_FILE  # reverse-sweep kernels are several times larger
    return [head + "".join(kernels[i : i + per]) for i in range(0, len(kernels), per)]


def _digest(sources: list[str]) -> str:
    h = hashlib.sha1()
    headers = ("tile.hpp", "common.hpp", "adjoint.hpp")
    for part in [" ".join(_build.COMMON_FLAGS)] + [(CSRC / f).read_text() for f in headers] + sources:
        h.update(part.encode())
    return h.hexdigest()[:20]


def _compile(sources: list[str], out: Path, jobs: int | None = None) -> None:
    work = out.with_suffix(".d")
    work.mkdir(parents=True, exist_ok=True)

    def one(i_src):
        i, src = i_src
        path = work / f"part{i}.hip"
        path.write_text(src)
        obj = work / f"part{i}.o"
        cmd = [_build.HIPCC, *_build.COMMON_FLAGS, "-c", str(path), "-o", str(obj)]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            err = "\n".join(ln for ln in res.stderr.splitlines() if "packed-fp32-ops" not in ln)
            raise RuntimeError(f"hipcc failed on a specialised pass file:\n{err[-4000:]}")
        return obj

    jobs = jobs or max(1, min(len(sources), (os.cpu_count() or 2), 16))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(one, enumerate(sources)))
    tmp = out.with_suffix(".so.tmp")
    cmd = [_build.HIPCC, f"--offload-arch={_build.ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link of specialised passes failed:\n{res.stderr[-4000:]}")
    os.replace(tmp, out)


class SpecializedPasses:
    """The launch handles of a program's compiled passes (kept alive with their shared object)."""

    def __init__(self, path: Path, n_passes: int, prefix: str = "k_spec_"):
        self.path = path
        self.lib = ctypes.CDLL(str(path))
        handles = [ctypes.cast(getattr(self.lib, f"{prefix}{i}"), ctypes.c_void_p).value for i in range(n_passes)]
        self.handles = (ctypes.c_void_p * max(n_passes, 1))(*handles)
        self.n_passes = n_passes


_loaded: dict[str, SpecializedPasses] = {}


def specialize(packed, cache_dir: Path | None = None, adjoint: bool = False) -> SpecializedPasses:
    """Compiles (or finds in the cache) the straight-line kernels of a PackedProgram: forward
    passes for cqs_run_passes_specialized, or (adjoint=True) reverse-sweep passes for
    cqs_adjoint_passes_specialized."""
    sources = program_sources(packed, adjoint)
    key = _digest(sources)
    if key in _loaded:
        return _loaded[key]
    cache = Path(cache_dir) if cache_dir is not None else CACHE_DIR
    cache.mkdir(parents=True, exist_ok=True)
    out = cache / f"passes_{key}.so"
    if not out.exists():
        _compile(sources, out)
    sp = SpecializedPasses(out, int(packed.n_passes), "k_aspec_" if adjoint else "k_spec_")
    _loaded[key] = sp
    return sp
