"""Serialized gate lists and deterministic circuit generators.

A circuit is a header plus one operation per line, named after the
``iqs::QubitRegister`` method it calls (include/qureg.hpp:136-372), e.g.::

    iqs-circuit 1
    qubits 20
    precision dp
    init base 0
    begin
    ApplyHadamard 0
    ApplyRotationX 3 1.2345
    ApplyCPauliX 0 1
    ZAll

Complex arguments are written as ``re im`` pairs (row-major for matrices); floats use
``repr`` so they round-trip bit-exactly.  Read-out lines (``GetProbability``,
``ExpectationValue``, ``ComputeNorm``, ``GetGlobalAmplitude``, ``ZAll``,
``QaoaExpectation``) produce ``R ...`` result lines.

Random numbers come from SplitMix64 (Steele, Lea, Flood 2014), so the same seed gives the
same gate list on every host.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Sequence, Tuple

import numpy as np

MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 generator; ``uniform()`` returns the top 53 bits scaled to [0, 1)."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, a: float = 0.0, b: float = 1.0) -> float:
        return a + (b - a) * ((self.next_u64() >> 11) * (1.0 / 9007199254740992.0))

    def randint(self, n: int) -> int:
        """Uniform integer in [0, n) by rejection (no modulo bias)."""
        lim = (1 << 64) - ((1 << 64) % n)
        while True:
            x = self.next_u64()
            if x < lim:
                return x % n

    def normal(self) -> float:
        """Standard normal by Box-Muller on two uniforms."""
        u1 = self.uniform()
        while u1 <= 0.0:
            u1 = self.uniform()
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


@dataclass
class Circuit:
    num_qubits: int
    precision: str = "dp"  # "dp" (ComplexDP) | "sp" (ComplexSP)
    init: str = "base"  # "base" | "plus"
    base_index: int = 0
    graph: List[Tuple[int, int, float]] = field(default_factory=list)
    ops: List[Tuple] = field(default_factory=list)

    def add(self, name: str, *args) -> "Circuit":
        self.ops.append((name,) + tuple(args))
        return self

    # ---- serialization ----
    @staticmethod
    def _fmt(a) -> str:
        if isinstance(a, (bool, np.bool_)):
            return "1" if a else "0"
        if isinstance(a, (int, np.integer)):
            return str(int(a))
        if isinstance(a, (complex, np.complexfloating)):
            return f"{float(a.real)!r} {float(a.imag)!r}"
        if isinstance(a, (float, np.floating)):
            return repr(float(a))
        if isinstance(a, np.ndarray):
            return " ".join(Circuit._fmt(x) for x in a.reshape(-1))
        if isinstance(a, (list, tuple)):
            return " ".join(Circuit._fmt(x) for x in a)
        return str(a)

    def to_text(self) -> str:
        lines = ["iqs-circuit 1", f"qubits {self.num_qubits}", f"precision {self.precision}"]
        lines.append("init plus" if self.init == "plus" else f"init base {self.base_index}")
        if self.graph:
            g = " ".join(f"{u} {v} {float(w)!r}" for (u, v, w) in self.graph)
            lines.append(f"graph {self.num_qubits} {len(self.graph)} {g}")
        lines.append("begin")
        for op in self.ops:
            lines.append(" ".join([op[0]] + [self._fmt(a) for a in op[1:]]))
        return "\n".join(lines) + "\n"

    def write(self, path: str) -> None:
        with open(path, "w") as f:
            f.write(self.to_text())

    @staticmethod
    def parse(text: str) -> "Circuit":
        c = Circuit(0)
        it = iter(text.splitlines())
        for line in it:
            t = line.split()
            if not t:
                continue
            if t[0] == "begin":
                break
            if t[0] == "qubits":
                c.num_qubits = int(t[1])
            elif t[0] == "precision":
                c.precision = t[1]
            elif t[0] == "init":
                c.init = t[1]
                if t[1] == "base":
                    c.base_index = int(t[2])
            elif t[0] == "graph":
                ne = int(t[2])
                vals = t[3:]
                c.graph = [(int(vals[3 * e]), int(vals[3 * e + 1]), float(vals[3 * e + 2])) for e in range(ne)]
        for line in it:
            t = line.split()
            if not t or t[0].startswith("#"):
                continue
            c.ops.append(_parse_op(t))
        return c


_INT_OPS_1 = {"ApplyPauliX", "ApplyPauliY", "ApplyPauliZ", "ApplyPauliSqrtX", "ApplyPauliSqrtY",
              "ApplyPauliSqrtZ", "ApplyT", "ApplyHadamard", "GetProbability"}
_INT_OPS_2 = {"ApplyCPauliX", "ApplyCPauliY", "ApplyCPauliZ", "ApplyCPauliSqrtZ", "ApplyCHadamard",
              "ApplySwap", "ApplyISwap", "ApplySqrtISwap", "Apply4thRootISwap", "EmulateSwap",
              "CollapseQubit"}
_ANGLE_1 = {"ApplyRotationX", "ApplyRotationY", "ApplyRotationZ"}
_ANGLE_2 = {"ApplyCRotationX", "ApplyCRotationY", "ApplyCRotationZ", "ApplyCPhaseRotation"}


def _cvec(vals: Sequence[str]) -> np.ndarray:
    v = np.array([float(x) for x in vals], dtype=np.float64)
    return v[0::2] + 1j * v[1::2]


def _parse_op(t: List[str]) -> Tuple:
    name = t[0]
    if name in _INT_OPS_1:
        return (name, int(t[1]))
    if name in _INT_OPS_2:
        return (name, int(t[1]), int(t[2]))
    if name in _ANGLE_1:
        return (name, int(t[1]), float(t[2]))
    if name in _ANGLE_2:
        return (name, int(t[1]), int(t[2]), float(t[3]))
    if name == "ApplyRotationXY":
        return (name, int(t[1]), float(t[2]), float(t[3]))
    if name == "Apply1QubitGate":
        return (name, int(t[1]), _cvec(t[2:10]).reshape(2, 2))
    if name in ("ApplyControlled1QubitGate", "ApplyISwapRotation"):
        return (name, int(t[1]), int(t[2]), _cvec(t[3:11]).reshape(2, 2))
    if name == "ApplyDiag":
        return (name, int(t[1]), int(t[2]), _cvec(t[3:11]))
    if name == "Apply2QubitGate":
        return (name, int(t[1]), int(t[2]), _cvec(t[3:35]).reshape(4, 4))
    if name == "ApplyToffoli":
        return (name, int(t[1]), int(t[2]), int(t[3]))
    if name == "PermuteQubits":
        return (name, t[1], [int(x) for x in t[2:]])
    if name == "ExpectationValue":
        k = int(t[2])
        return (name, float(t[1]), k, [int(x) for x in t[3:3 + 2 * k]])
    if name == "QaoaPhase":
        return (name, float(t[1]))
    if name in ("Print", "ExportAmplitudes", "dumpbin", "loadbin"):
        return (name, t[1])
    if name == "SetRngSeed":
        return (name, int(t[1]))
    if name == "SetNoiseTimescales":
        return (name, float(t[1]), float(t[2]))
    if name == "ApplyNoiseGate":
        return (name, int(t[1]), float(t[2]))
    if name == "TurnOnFusion":
        return (name, int(t[1]))
    if name == "GetGlobalAmplitude":
        return (name, int(t[1]))
    return (name,)


def format_expectation(coeff: float, pairs: Sequence[Tuple[int, int]]) -> Tuple:
    """Build an ExpectationValue op: pairs of (qubit, observable 1|2|3 = X|Y|Z)."""
    return ("ExpectationValue", float(coeff), len(pairs), [x for p in pairs for x in p])


# ---------------------------------------------------------------------------------------
# Random matrices


def random_su2(rng: SplitMix64) -> np.ndarray:
    """Haar-random SU(2) from a uniformly random unit quaternion."""
    q = np.array([rng.normal() for _ in range(4)])
    q /= np.linalg.norm(q)
    a = q[0] + 1j * q[1]
    b = q[2] + 1j * q[3]
    return np.array([[a, -np.conj(b)], [b, np.conj(a)]], dtype=np.complex128)


def haar_unitary(d: int, rng: SplitMix64) -> np.ndarray:
    """Haar-random d x d unitary: QR of a complex Gaussian matrix, R's diagonal phase-fixed
    (Mezzadri 2007)."""
    z = np.array([[complex(rng.normal(), rng.normal()) for _ in range(d)] for _ in range(d)]) / math.sqrt(2.0)
    q, r = np.linalg.qr(z)
    ph = np.diag(r) / np.abs(np.diag(r))
    return q * ph[np.newaxis, :]


# ---------------------------------------------------------------------------------------
# Workload generators


def brick_circuit(n: int, depth: int, seed: int, precision: str = "dp", readout: bool = True) -> Circuit:
    """Random brick circuit from |0...0>: each layer applies H on every qubit, then
    R{X,Y,Z}(theta) on every qubit (axis uniform, theta ~ U[0, 2pi)), then CNOT(i, i+1) for
    i = layer mod 2, layer mod 2 + 2, ...  Read-out: <Z_i> for every i."""
    rng = SplitMix64(seed)
    c = Circuit(n, precision=precision)
    axes = ("ApplyRotationX", "ApplyRotationY", "ApplyRotationZ")
    for layer in range(depth):
        for q in range(n):
            c.add("ApplyHadamard", q)
        for q in range(n):
            ax = axes[rng.randint(3)]
            c.add(ax, q, rng.uniform(0.0, 2.0 * math.pi))
        for q in range(layer % 2, n - 1, 2):
            c.add("ApplyCPauliX", q, q + 1)
    if readout:
        c.add("ZAll")
    return c


def brick_layer_ops(n: int, layer: int, rng: SplitMix64) -> List[Tuple]:
    """The ops of one brick layer (same draw order as brick_circuit)."""
    axes = ("ApplyRotationX", "ApplyRotationY", "ApplyRotationZ")
    ops: List[Tuple] = [("ApplyHadamard", q) for q in range(n)]
    for q in range(n):
        ax = axes[rng.randint(3)]
        ops.append((ax, q, rng.uniform(0.0, 2.0 * math.pi)))
    ops += [("ApplyCPauliX", q, q + 1) for q in range(layer % 2, n - 1, 2)]
    return ops


def random_1q2q_circuit(n: int, num_gates: int, seed: int, precision: str = "dp") -> Circuit:
    """num_gates Haar-random gates: each is 1q (probability 1/2, target uniform) or a dense
    2q gate on two distinct uniform qubits, starting from |0...0>."""
    rng = SplitMix64(seed)
    c = Circuit(n, precision=precision)
    for _ in range(num_gates):
        if rng.uniform() < 0.5:
            q = rng.randint(n)
            c.add("Apply1QubitGate", q, haar_unitary(2, rng))
        else:
            a = rng.randint(n)
            b = rng.randint(n - 1)
            b = b + 1 if b >= a else b
            c.add("Apply2QubitGate", a, b, haar_unitary(4, rng))
    return c


def random_regular_graph(nv: int, degree: int, seed: int) -> List[Tuple[int, int, float]]:
    """Random d-regular simple graph by configuration-model rejection sampling, unit weights."""
    rng = SplitMix64(seed)
    while True:
        stubs = [v for v in range(nv) for _ in range(degree)]
        for i in range(len(stubs) - 1, 0, -1):  # Fisher-Yates
            j = rng.randint(i + 1)
            stubs[i], stubs[j] = stubs[j], stubs[i]
        edges = set()
        ok = True
        for k in range(0, len(stubs), 2):
            u, v = stubs[k], stubs[k + 1]
            if u == v or (min(u, v), max(u, v)) in edges:
                ok = False
                break
            edges.add((min(u, v), max(u, v)))
        if ok:
            return [(u, v, 1.0) for (u, v) in sorted(edges)]


def qaoa_circuit(n: int, graph, gammas: Sequence[float], betas: Sequence[float],
                 precision: str = "dp") -> Circuit:
    """QAOA MaxCut from |+>^n: per layer exp(-i gamma C) (ImplementQaoaLayerBasedOnCostFunction)
    then exp(-i beta sum X) = prod_q RX_q(2 beta) in the RX(theta) = exp(-i theta X / 2)
    convention of ApplyRotationX.  Read-out: <C>."""
    c = Circuit(n, precision=precision, init="plus", graph=list(graph))
    for g, b in zip(gammas, betas):
        c.add("QaoaPhase", float(g))
        for q in range(n):
            c.add("ApplyRotationX", q, 2.0 * float(b))
    c.add("QaoaExpectation")
    return c


def hea_instances(count: int, seed: int, nmin: int = 10, nmax: int = 22,
                  dmin: int = 10, dmax: int = 50):
    """Hardware-efficient ansatz instances: (n, depth, thetas[depth, n], phis[depth, n]).
    Each layer is RY(theta) then RZ(phi) on every qubit, then CZ(i, i+1) for all i."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        n = nmin + rng.randint(nmax - nmin + 1)
        d = dmin + rng.randint(dmax - dmin + 1)
        th = np.array([[rng.uniform(0.0, 2 * math.pi) for _ in range(n)] for _ in range(d)])
        ph = np.array([[rng.uniform(0.0, 2 * math.pi) for _ in range(n)] for _ in range(d)])
        out.append((n, d, th, ph))
    return out
