"""Cirq's JSON wire format for the objects this simulator consumes.

`to_json` / `read_json` follow cirq-core/cirq/protocols/json_serialization.py (to_json at :468-519,
read_json, the `cirq_type` tag and `_json_dict_` field names of each class), so that circuits,
observables, resolvers and sweeps written by any Cirq installation load here, and the other way
round.  Only simulable qubit-based types are covered: qubits (LineQubit, GridQubit, NamedQubit),
Circuit / Moment / GateOperation / ControlledOperation / Pauli-string operations, the gate and
channel families of cirq_amd.gates and cirq_amd.channels, PauliString / PauliSum, ParamResolver,
Points / Linspace / Zip / Product sweeps, complex numbers, numpy scalars and the sympy expression
nodes (Symbol, Add, Mul, Pow, Integer, Float, Rational, pi).  Anything else raises ValueError.
"""

from __future__ import annotations

import json
import numbers
import pathlib

import numpy as np
import sympy

from . import channels as ch
from . import gates as g
from . import paulis as pl
from .circuits import Circuit, Moment
from .qubits import GridQubit, LineQubit, NamedQubit
from .study import Linspace, ParamResolver, Points, Product, Zip

_EIGEN = {
    "XPowGate": g.XPowGate, "YPowGate": g.YPowGate, "ZPowGate": g.ZPowGate, "HPowGate": g.HPowGate,
    "CZPowGate": g.CZPowGate, "CXPowGate": g.CXPowGate, "CNotPowGate": g.CXPowGate, "SwapPowGate": g.SwapPowGate,
    "ISwapPowGate": g.ISwapPowGate, "XXPowGate": g.XXPowGate, "YYPowGate": g.YYPowGate, "ZZPowGate": g.ZZPowGate,
    "CCZPowGate": g.CCZPowGate, "CCXPowGate": g.CCXPowGate,
}
_ROT = {"Rx": g.rx, "Ry": g.ry, "Rz": g.rz}
_PAULI = {"_PauliX": pl.X, "_PauliY": pl.Y, "_PauliZ": pl.Z}


# ------------------------------------------------------------------------------ decoding
def _complex(d):
    return complex(d["real"], d["imag"])


def _num(v):
    return _decode(v)


def _qubits(lst):
    return [_decode(q) for q in lst]


def _eigen(cls):
    def make(d):
        if d.get("dimension", 2) != 2:
            raise ValueError(f"{d['cirq_type']} with dimension {d['dimension']}: only qubits are supported")
        return cls(exponent=_num(d["exponent"]), global_shift=float(d.get("global_shift", 0.0)))

    return make


def _pauli(name):
    def make(d):
        p = _PAULI[name]
        e = _num(d.get("exponent", 1))
        return p if e == 1 else p**e

    return make


def _controlled_values(cv):
    """Control values: a list of ints, or a ProductOfSums with one value per control."""
    if isinstance(cv, dict):
        if cv.get("cirq_type") != "ProductOfSums":
            raise ValueError(f"control values {cv.get('cirq_type')} are not supported")
        cv = cv["data"]
    out = []
    for v in cv:
        vals = v if isinstance(v, (list, tuple)) else [v]
        if len(vals) != 1 or vals[0] not in (0, 1):
            raise ValueError("only one control value (0 or 1) per qubit is supported")
        out.append(int(vals[0]))
    return out


def _controlled_gate(d):
    shape = d.get("control_qid_shape")
    if shape is not None and any(s != 2 for s in shape):
        raise ValueError("controls must be qubits")
    cv = _controlled_values(d["control_values"])
    return g.ControlledGate(_decode(d["sub_gate"]), num_controls=len(cv), control_values=cv)


def _controlled_operation(d):
    controls = _qubits(d["controls"])
    cv = _controlled_values(d["control_values"])
    sub = _decode(d["sub_operation"])
    gate = getattr(sub, "gate", None)
    if gate is None:
        raise ValueError("controlled operations need a gate operation inside")
    return g.ControlledGate(gate, num_controls=len(cv), control_values=cv).on(*controls, *sub.qubits)


def _measurement(d):
    if d.get("qid_shape") is not None and any(s != 2 for s in d["qid_shape"]):
        raise ValueError("measurements of qudits are not supported")
    if d.get("confusion_map"):
        raise ValueError("measurement confusion maps are not supported")
    key = d["key"]
    if isinstance(key, dict):
        key = key.get("name", "")
    return g.MeasurementGate(d["num_qubits"], key, tuple(d.get("invert_mask", ())))


def _matrix_gate(d):
    shape = d.get("qid_shape")
    if shape is not None and any(s != 2 for s in shape):
        raise ValueError("MatrixGate on qudits is not supported")
    m = np.array([[_decode(x) for x in row] for row in d["matrix"]], dtype=np.complex128)
    return g.MatrixGate(m, name=d.get("name"))


def _pauli_string(d):
    m = {_decode(q): _decode(p) for q, p in d["qubit_pauli_map"]}
    return pl.PauliString(qubit_pauli_map=m, coefficient=_decode(d.get("coefficient", 1)))


def _pauli_sum(d):
    terms = []
    for qmap, coef in d["items"]:
        m = {_decode(q): _decode(p) for q, p in qmap}
        terms.append(pl.PauliString(qubit_pauli_map=m, coefficient=_decode(coef)))
    return pl.PauliSum.from_pauli_strings(terms)


def _param_resolver(d):
    out = {}
    for k, v in d["param_dict"]:
        k = _decode(k)
        out[k.name if isinstance(k, sympy.Symbol) else k] = _decode(v)
    return ParamResolver(out)


def _sympy_args(d):
    return [_decode(a) for a in d["args"]]


_DECODERS = {
    "LineQubit": lambda d: LineQubit(d["x"]),
    "GridQubit": lambda d: GridQubit(d["row"], d["col"]),
    "NamedQubit": lambda d: NamedQubit(d["name"]),
    "complex": _complex,
    "Circuit": lambda d: Circuit.from_moments(*(_decode(m) for m in d["moments"])),
    "Moment": lambda d: Moment([_decode(o) for o in d["operations"]]),
    "GateOperation": lambda d: _decode(d["gate"]).on(*_qubits(d["qubits"])),
    "SingleQubitPauliStringGateOperation": lambda d: _decode(d["pauli"]).on(_decode(d["qubit"])),
    "ControlledOperation": _controlled_operation,
    "ControlledGate": _controlled_gate,
    "PhasedXPowGate": lambda d: g.PhasedXPowGate(
        phase_exponent=_num(d["phase_exponent"]), exponent=_num(d.get("exponent", 1.0)),
        global_shift=float(d.get("global_shift", 0.0))),
    "PhasedXZGate": lambda d: g.PhasedXZGate(
        x_exponent=_num(d["x_exponent"]), z_exponent=_num(d["z_exponent"]),
        axis_phase_exponent=_num(d["axis_phase_exponent"])),
    "FSimGate": lambda d: g.FSimGate(theta=_num(d["theta"]), phi=_num(d["phi"])),
    "MatrixGate": _matrix_gate,
    "DiagonalGate": lambda d: g.DiagonalGate([_num(a) for a in d["diag_angles_radians"]]),
    "IdentityGate": lambda d: (g.IdentityGate(d["num_qubits"]) if not d.get("qid_shape")
                               or all(s == 2 for s in d["qid_shape"]) else _unsupported("IdentityGate on qudits")),
    "GlobalPhaseGate": lambda d: g.GlobalPhaseGate(_decode(d["coefficient"])),
    "CSwapGate": lambda d: g.CSWAP if hasattr(g, "CSWAP") else g.CSwapGate(),
    "MeasurementGate": _measurement,
    "DepolarizingChannel": lambda d: ch.DepolarizingChannel(d["p"], d.get("n_qubits", 1)),
    "AsymmetricDepolarizingChannel": lambda d: ch.AsymmetricDepolarizingChannel(
        error_probabilities=d["error_probabilities"]),
    "BitFlipChannel": lambda d: ch.BitFlipChannel(d["p"]),
    "PhaseFlipChannel": lambda d: ch.PhaseFlipChannel(d["p"]),
    "AmplitudeDampingChannel": lambda d: ch.AmplitudeDampingChannel(d["gamma"]),
    "GeneralizedAmplitudeDampingChannel": lambda d: ch.GeneralizedAmplitudeDampingChannel(d["p"], d["gamma"]),
    "PhaseDampingChannel": lambda d: ch.PhaseDampingChannel(d["gamma"]),
    "ResetChannel": lambda d: (ch.ResetChannel() if d.get("dimension", 2) == 2
                               else _unsupported("ResetChannel on qudits")),
    "PauliString": _pauli_string,
    "PauliSum": _pauli_sum,
    "ParamResolver": _param_resolver,
    "Points": lambda d: _with_metadata(Points(d["key"], [_decode(p) for p in d["points"]]), d),
    "Linspace": lambda d: _with_metadata(Linspace(d["key"], d["start"], d["stop"], d["length"]), d),
    "Zip": lambda d: Zip(*(_decode(s) for s in d["sweeps"])),
    "Product": lambda d: Product(*(_decode(s) for s in d["factors"])),
    "sympy.Symbol": lambda d: sympy.Symbol(d["name"]),
    "sympy.Add": lambda d: sympy.Add(*_sympy_args(d)),
    "sympy.Mul": lambda d: sympy.Mul(*_sympy_args(d)),
    "sympy.Pow": lambda d: sympy.Pow(*_sympy_args(d)),
    "sympy.Integer": lambda d: sympy.Integer(d["i"]),
    "sympy.Float": lambda d: sympy.Float(d["approx"]),
    "sympy.Rational": lambda d: sympy.Rational(d["p"], d["q"]),
    "sympy.pi": lambda d: sympy.pi,
}
_DECODERS.update({k: _eigen(v) for k, v in _EIGEN.items()})
_DECODERS.update({k: (lambda f: (lambda d: f(_num(d["rads"]))))(v) for k, v in _ROT.items()})
_DECODERS.update({k: _pauli(k) for k in _PAULI})


def _with_metadata(sweep, d):
    if "metadata" in d:
        sweep.metadata = _decode(d["metadata"])
    return sweep


def _unsupported(what):
    raise ValueError(f"{what} is not supported")


def _decode(v):
    if isinstance(v, list):
        return [_decode(x) for x in v]
    if not isinstance(v, dict):
        return v
    t = v.get("cirq_type")
    if t is None:
        return {k: _decode(x) for k, x in v.items()}
    f = _DECODERS.get(t)
    if f is None:
        raise ValueError(f"cirq_type {t!r} is not supported by cirq_amd")
    return f(v)


def read_json(file_or_fn=None, *, json_text: str | None = None):
    """Object(s) from Cirq JSON text, a path or an open file (cirq.read_json's arguments)."""
    if (file_or_fn is None) == (json_text is None):
        raise ValueError("Must specify ONE of 'file_or_fn' or 'json_text'.")
    if json_text is None:
        if isinstance(file_or_fn, (str, pathlib.Path)):
            with open(file_or_fn) as f:
                json_text = f.read()
        else:
            json_text = file_or_fn.read()
    return _decode(json.loads(json_text))


# ------------------------------------------------------------------------------ encoding
def _enc_sympy(e):
    if isinstance(e, sympy.Symbol):
        return {"cirq_type": "sympy.Symbol", "name": e.name}
    if e is sympy.pi:
        return {"cirq_type": "sympy.pi"}
    if isinstance(e, sympy.Integer):
        return {"cirq_type": "sympy.Integer", "i": int(e)}
    if isinstance(e, sympy.Rational):
        return {"cirq_type": "sympy.Rational", "p": int(e.p), "q": int(e.q)}
    if isinstance(e, sympy.Float):
        return {"cirq_type": "sympy.Float", "approx": float(e)}
    for cls, name in ((sympy.Add, "sympy.Add"), (sympy.Mul, "sympy.Mul"), (sympy.Pow, "sympy.Pow")):
        if isinstance(e, cls):
            return {"cirq_type": name, "args": [_encode(a) for a in e.args]}
    raise ValueError(f"sympy expression {e!r} cannot be written as Cirq JSON")


def _enc_gate(gt):
    name = type(gt).__name__
    if name in _PAULI:
        return {"cirq_type": name, "exponent": _encode(gt.exponent), "global_shift": float(gt.global_shift)}
    if name in _ROT:
        return {"cirq_type": name, "rads": _encode(gt.rads)}
    if name in _EIGEN:
        return {"cirq_type": name, "exponent": _encode(gt.exponent), "global_shift": float(gt.global_shift)}
    if name == "PhasedXPowGate":
        return {"cirq_type": name, "phase_exponent": _encode(gt.phase_exponent), "exponent": _encode(gt.exponent),
                "global_shift": float(gt.global_shift)}
    if name == "PhasedXZGate":
        return {"cirq_type": name, "axis_phase_exponent": _encode(gt.axis_phase_exponent),
                "x_exponent": _encode(gt.x_exponent), "z_exponent": _encode(gt.z_exponent)}
    if name == "FSimGate":
        return {"cirq_type": name, "theta": _encode(gt.theta), "phi": _encode(gt.phi)}
    if name == "MatrixGate":
        d = {"cirq_type": name, "matrix": _encode(np.asarray(gt._unitary_()))}
        if getattr(gt, "_name", None):
            d["name"] = gt._name
        return d
    if name == "DiagonalGate":
        return {"cirq_type": name, "diag_angles_radians": [_encode(a) for a in gt.diag_angles_radians]}
    if name == "IdentityGate":
        return {"cirq_type": name, "num_qubits": gt.num_qubits()}
    if name == "GlobalPhaseGate":
        return {"cirq_type": name, "coefficient": _encode(gt.coefficient)}
    if name == "CSwapGate":
        return {"cirq_type": name}
    if name == "MeasurementGate":
        return {"cirq_type": name, "num_qubits": gt.num_qubits(), "key": gt.key,
                "invert_mask": list(gt.invert_mask)}
    if name == "ControlledGate":
        return {"cirq_type": name, "control_values": {"cirq_type": "ProductOfSums",
                                                      "data": [[int(v)] for v in gt.control_values]},
                "control_qid_shape": [2] * gt.num_controls, "sub_gate": _encode(gt.sub_gate)}
    if name == "DepolarizingChannel":
        d = {"cirq_type": name, "p": gt.p}
        if gt.n_qubits != 1:
            d["n_qubits"] = gt.n_qubits
        return d
    if name == "AsymmetricDepolarizingChannel":
        return {"cirq_type": name, "error_probabilities": dict(gt.error_probabilities)}
    if name in ("BitFlipChannel", "PhaseFlipChannel"):
        return {"cirq_type": name, "p": gt.p}
    if name in ("AmplitudeDampingChannel", "PhaseDampingChannel"):
        return {"cirq_type": name, "gamma": gt.gamma}
    if name == "GeneralizedAmplitudeDampingChannel":
        return {"cirq_type": name, "p": gt.p, "gamma": gt.gamma}
    if name == "ResetChannel":
        return {"cirq_type": name, "dimension": 2}
    raise ValueError(f"{gt!r} cannot be written as Cirq JSON")


def _encode(o):
    if isinstance(o, sympy.Basic):
        return _enc_sympy(o)
    if isinstance(o, (bool, np.bool_)):
        return bool(o)
    if isinstance(o, (int, np.integer)):
        return int(o)
    if isinstance(o, (float, np.floating)):
        return float(o)
    if isinstance(o, (complex, np.complexfloating)):
        return {"cirq_type": "complex", "real": float(o.real), "imag": float(o.imag)}
    if isinstance(o, np.ndarray):
        return [_encode(x) for x in o.tolist()] if o.ndim else _encode(o.item())
    if isinstance(o, (list, tuple)):
        return [_encode(x) for x in o]
    if isinstance(o, str) or o is None:
        return o
    if isinstance(o, LineQubit):
        return {"cirq_type": "LineQubit", "x": o.x}
    if isinstance(o, GridQubit):
        return {"cirq_type": "GridQubit", "row": o.row, "col": o.col}
    if isinstance(o, NamedQubit):
        return {"cirq_type": "NamedQubit", "name": o.name}
    if isinstance(o, Circuit):
        return {"cirq_type": "Circuit", "moments": [_encode(m) for m in o]}
    if isinstance(o, Moment):
        return {"cirq_type": "Moment", "operations": [_encode(op) for op in o.operations]}
    if isinstance(o, pl.PauliSum):
        return {"cirq_type": "PauliSum", "items": [
            [[[_encode(q), _encode(p)] for q, p in t.items()], _encode(complex(t.coefficient))] for t in o]}
    if isinstance(o, pl.SingleQubitPauliStringGateOperation) and complex(o.coefficient) == 1:
        (q, p), = o.items()
        return {"cirq_type": "SingleQubitPauliStringGateOperation", "pauli": _encode(p), "qubit": _encode(q)}
    if isinstance(o, pl.PauliString):
        return {"cirq_type": "PauliString", "qubit_pauli_map": [[_encode(q), _encode(p)] for q, p in o.items()],
                "coefficient": _encode(complex(o.coefficient))}
    if isinstance(o, g.GateOperation):
        return {"cirq_type": "GateOperation", "gate": _encode(o.gate), "qubits": [_encode(q) for q in o.qubits]}
    if isinstance(o, g.Gate):
        return _enc_gate(o)
    if isinstance(o, ParamResolver):
        return {"cirq_type": "ParamResolver", "param_dict": [[_encode(k), _encode(v)]
                                                             for k, v in o.param_dict.items()]}
    if isinstance(o, Points):
        return _add_metadata({"cirq_type": "Points", "key": o.key, "points": _encode(list(o.points))}, o)
    if isinstance(o, Linspace):
        return _add_metadata({"cirq_type": "Linspace", "key": o.key, "start": o.start, "stop": o.stop,
                              "length": o.length}, o)
    if isinstance(o, Zip):
        return {"cirq_type": "Zip", "sweeps": [_encode(s) for s in o.sweeps]}
    if isinstance(o, Product):
        return {"cirq_type": "Product", "factors": [_encode(s) for s in o.factors]}
    if isinstance(o, dict):
        return {k: _encode(v) for k, v in o.items()}
    if isinstance(o, numbers.Number):
        return complex(o)
    raise ValueError(f"{o!r} cannot be written as Cirq JSON")


def _add_metadata(d, sweep):
    if getattr(sweep, "metadata", None) is not None:
        d["metadata"] = _encode(sweep.metadata)
    return d


def to_json(obj, file_or_fn=None, *, indent: int | None = 2, separators=None) -> str | None:
    """Cirq JSON text of obj; written to file_or_fn (path or open file) when given."""
    text = json.dumps(_encode(obj), indent=indent, separators=separators)
    if file_or_fn is None:
        return text
    if isinstance(file_or_fn, (str, pathlib.Path)):
        with open(file_or_fn, "w") as f:
            f.write(text)
    else:
        file_or_fn.write(text)
    return None
