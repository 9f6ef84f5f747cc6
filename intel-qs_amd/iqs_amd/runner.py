"""Execute a serialized circuit (iqs_amd.circuits format) on the intelqs_py drop-in.

Each line calls the QubitRegister method of the same name, so a gate list written for the
reference runs unchanged on the GPU register.  Read-out lines return "R <op> <values>"
strings formatted like the reference runner's (%.17g).
"""
from __future__ import annotations

import os
import sys
from typing import List, Tuple

import numpy as np

from .circuits import Circuit

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if _PKG not in sys.path:
    sys.path.insert(0, _PKG)


def _module():
    import intelqs_py  # the drop-in module, built in-tree

    return intelqs_py


def _fmt(x: float) -> str:
    return "%.17g" % x


def new_register(c: Circuit):
    m = _module()
    cls = m.QubitRegisterSP if c.precision == "sp" else m.QubitRegister
    if c.init == "plus":
        return cls(c.num_qubits, "++++", 0, 0)
    return cls(c.num_qubits, "base", c.base_index, 0)


def run(c: Circuit, reg=None) -> Tuple[object, List[str]]:
    m = _module()
    cdt = np.complex64 if c.precision == "sp" else np.complex128
    reg = new_register(c) if reg is None else reg
    out: List[str] = []
    keep: List[object] = []  # objects the register points to (generators)
    diag = None

    def ensure_diag():
        nonlocal diag
        if diag is None:
            n = c.num_qubits
            diag = (m.QubitRegisterSP if c.precision == "sp" else m.QubitRegister)(n, "base", 0, 0)
            weighted = any(w != 1.0 for (_, _, w) in c.graph)
            if weighted:
                adj = [0.0] * (n * n)
                for (u, v, w) in c.graph:
                    adj[u * n + v] = adj[v * n + u] = float(w)
                m.InitializeVectorAsWeightedMaxCutCostFunction(diag, adj)
            else:
                adj = [0] * (n * n)
                for (u, v, _) in c.graph:
                    adj[u * n + v] = adj[v * n + u] = 1
                m.InitializeVectorAsMaxCutCostFunction(diag, adj)
        return diag

    for op in c.ops:
        name, args = op[0], op[1:]
        if name in ("Apply1QubitGate",):
            reg.Apply1QubitGate(args[0], np.asarray(args[1], dtype=cdt))
        elif name in ("ApplyControlled1QubitGate", "ApplyISwapRotation"):
            getattr(reg, name)(args[0], args[1], np.asarray(args[2], dtype=cdt))
        elif name == "ApplyDiag":
            reg.ApplyDiag(args[0], args[1], np.diag(np.asarray(args[2])).astype(cdt))
        elif name == "Apply2QubitGate":
            reg.Apply2QubitGate(args[0], args[1], np.asarray(args[2], dtype=cdt))
        elif name == "PermuteQubits":
            reg.PermuteQubits(list(args[1]), args[0])
        elif name == "CollapseQubit":
            reg.CollapseQubit(args[0], bool(args[1]))
        elif name == "TurnOnFusion":
            reg.TurnOnFusion(args[0])
        elif name == "SetRngSeed":
            # the module binds RandomNumberGenerator for ComplexDP registers (as the reference's)
            rng = m.RandomNumberGenerator()
            rng.SetSeedStreamPtrs(args[0])
            reg.SetRngPtr(rng)
            keep.append(rng)
        elif name == "QaoaPhase":
            m.ImplementQaoaLayerBasedOnCostFunction(reg, ensure_diag(), args[0])
        elif name == "GetProbability":
            out.append(f"R GetProbability {args[0]} {_fmt(reg.GetProbability(args[0]))}")
        elif name == "ExpectationValue":
            coeff, k, flat = args
            qs = [flat[2 * i] for i in range(k)]
            os_ = [flat[2 * i + 1] for i in range(k)]
            out.append(f"R ExpectationValue {_fmt(reg.ExpectationValue(qs, os_, coeff))}")
        elif name == "ComputeNorm":
            out.append(f"R ComputeNorm {_fmt(reg.ComputeNorm())}")
        elif name == "Entropy":
            out.append(f"R Entropy {_fmt(reg.Entropy())}")
        elif name == "GoogleStats":
            out.append("R GoogleStats " + " ".join(_fmt(v) for v in reg.GoogleStats()))
        elif name == "GetGlobalAmplitude":
            v = reg.GetGlobalAmplitude(args[0])
            out.append(f"R GetGlobalAmplitude {args[0]} {_fmt(v.real)} {_fmt(v.imag)}")
        elif name == "QaoaExpectation":
            out.append(f"R QaoaExpectation {_fmt(m.GetExpectationValueFromCostFunction(reg, ensure_diag()))}")
        elif name == "ZAll":
            out.append("R ZAll " + " ".join(_fmt(reg.ExpectationValueZ(q)) for q in range(c.num_qubits)))
        else:
            getattr(reg, name)(*args)
    reg.TurnOffFusion()
    if keep:
        reg._iqs_keep = keep  # the register holds a raw pointer to the generator
    return reg, out


def state_of(reg) -> np.ndarray:
    """Final amplitudes in data order (a copy of the register's host mirror)."""
    return np.array(reg, copy=True)


def parse_results(lines: List[str]) -> List[Tuple[str, List[float]]]:
    res = []
    for l in lines:
        t = l.split()
        res.append((t[1], [float(x) for x in t[2:]]))
    return res
