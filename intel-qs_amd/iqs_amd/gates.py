"""Named-gate matrices of iqs::QubitRegister, evaluated exactly as the reference builds them
(double precision; a ComplexSP register narrows them to float when they cross the C ABI).

Sources: src/qureg_apply1qubitgate.cpp:277-501, src/qureg_applyctrl1qubitgate.cpp:458-618,
src/qureg_applyswap.cpp:23-108, src/qureg_applytoffoli.cpp:24-47.
"""
from __future__ import annotations

import cmath
import math

import numpy as np

SQ2 = 1.0 / math.sqrt(2.0)


def m2(a, b, c, d) -> np.ndarray:
    return np.array([[a, b], [c, d]], dtype=np.complex128)


def rx(theta: float) -> np.ndarray:
    c, s = math.cos(theta / 2.0), math.sin(theta / 2.0)
    return m2(c, complex(0, -s), complex(0, -s), c)


def ry(theta: float) -> np.ndarray:
    c, s = math.cos(theta / 2.0), math.sin(theta / 2.0)
    return m2(c, -s, s, c)


def rz(theta: float) -> np.ndarray:
    c, s = math.cos(theta / 2.0), math.sin(theta / 2.0)
    return m2(complex(c, -s), 0, 0, complex(c, s))


def rxy(phi: float, theta: float) -> np.ndarray:
    c, s = math.cos(theta / 2.0), math.sin(theta / 2.0)
    return m2(c, complex(-s * math.sin(phi), -s * math.cos(phi)), complex(s * math.sin(phi), -s * math.cos(phi)), c)


H = m2(SQ2, SQ2, SQ2, -SQ2)
X = m2(0, 1, 1, 0)
Y = m2(0, -1j, 1j, 0)
Z = m2(1, 0, 0, -1)
SQRT_X = m2(0.5 + 0.5j, 0.5 - 0.5j, 0.5 - 0.5j, 0.5 + 0.5j)
SQRT_Y = m2(0.5 + 0.5j, -0.5 - 0.5j, 0.5 + 0.5j, 0.5 + 0.5j)
SQRT_Z = m2(1, 0, 0, 1j)
T = m2(1, 0, 0, complex(math.cos(math.pi / 4.0), math.sin(math.pi / 4.0)))


def cphase(theta: float) -> np.ndarray:
    return m2(1, 0, 0, complex(math.cos(theta), math.sin(theta)))


# Swap family: 2x2 acting on (|01>, |10>)
SWAP = m2(0, 1, 1, 0)
ISWAP = m2(0, 1j, 1j, 0)
SQRT_ISWAP = m2(SQ2, 1j * SQ2, 1j * SQ2, SQ2)


def iswap_4th_root() -> np.ndarray:
    a = cmath.rect(0.5, math.pi / 8.0)
    b = cmath.rect(0.5, 7.0 * math.pi / 8.0)
    return m2(a - b, a + b, a + b, a - b)


# Toffoli decomposition factors
TOFFOLI_V = m2(0.5 - 0.5j, 0.5 + 0.5j, 0.5 + 0.5j, 0.5 - 0.5j)
TOFFOLI_VD = m2(0.5 + 0.5j, 0.5 - 0.5j, 0.5 - 0.5j, 0.5 + 0.5j)

NAMED_1Q = {
    "ApplyHadamard": lambda: H, "ApplyPauliX": lambda: X, "ApplyPauliY": lambda: Y,
    "ApplyPauliZ": lambda: Z, "ApplyPauliSqrtX": lambda: SQRT_X, "ApplyPauliSqrtY": lambda: SQRT_Y,
    "ApplyPauliSqrtZ": lambda: SQRT_Z, "ApplyT": lambda: T,
}
ROT_1Q = {"ApplyRotationX": rx, "ApplyRotationY": ry, "ApplyRotationZ": rz}
NAMED_C = {
    "ApplyCPauliX": lambda: X, "ApplyCPauliY": lambda: Y, "ApplyCPauliZ": lambda: Z,
    "ApplyCPauliSqrtZ": lambda: SQRT_Z, "ApplyCHadamard": lambda: H,
}
ROT_C = {"ApplyCRotationX": rx, "ApplyCRotationY": ry, "ApplyCRotationZ": rz, "ApplyCPhaseRotation": cphase}
