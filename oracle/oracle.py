"""ctypes front end of the CPU oracle (liboracle.so, built from iqs_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import List, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run `make -C oracle oracle`)")
        L = C.CDLL(path)
        L.orc_run.argtypes = [C.c_char_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]
        L.orc_run.restype = C.c_int
        L.orc_free.argtypes = [C.c_void_p]
        for sfx in ("dp", "sp"):
            P = C.c_void_p
            getattr(L, f"orc_1q_{sfx}").argtypes = [P, C.c_int, C.c_uint, P]
            getattr(L, f"orc_c1q_{sfx}").argtypes = [P, C.c_int, C.c_uint, C.c_uint, P]
            getattr(L, f"orc_swaplike_{sfx}").argtypes = [P, C.c_int, C.c_uint, C.c_uint, P]
            getattr(L, f"orc_diag_{sfx}").argtypes = [P, C.c_int, C.c_uint, C.c_uint, P]
            getattr(L, f"orc_2q_{sfx}").argtypes = [P, C.c_int, C.c_uint, C.c_uint, P]
            getattr(L, f"orc_permute_{sfx}").argtypes = [P, C.c_int, P, P]
            for f in ("prob1",):
                getattr(L, f"orc_{f}_{sfx}").argtypes = [P, C.c_int, C.c_uint]
                getattr(L, f"orc_{f}_{sfx}").restype = C.c_double
            getattr(L, f"orc_norm_{sfx}").argtypes = [P, C.c_int]
            getattr(L, f"orc_norm_{sfx}").restype = C.c_double
            getattr(L, f"orc_parity_norm_{sfx}").argtypes = [P, C.c_int, C.c_uint64]
            getattr(L, f"orc_parity_norm_{sfx}").restype = C.c_double
            getattr(L, f"orc_overlap_{sfx}").argtypes = [P, P, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
            getattr(L, f"orc_qaoa_expect_{sfx}").argtypes = [P, P, C.c_int]
            getattr(L, f"orc_qaoa_expect_{sfx}").restype = C.c_double
            getattr(L, f"orc_collapse_{sfx}").argtypes = [P, C.c_int, C.c_uint, C.c_int]
        L.orc_qaoa_phase_dp.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double]
        L.orc_qaoa_phase_sp.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_float]
        _LIB = L
    return _LIB


def run(text: str) -> Tuple[np.ndarray, List[str]]:
    """Run a serialized circuit; return (final state in data order, result lines)."""
    hdr = {}
    for line in text.splitlines():
        t = line.split()
        if not t:
            continue
        if t[0] == "begin":
            break
        hdr[t[0]] = t[1:]
    n = int(hdr["qubits"][0])
    dt = np.complex64 if hdr.get("precision", ["dp"])[0] == "sp" else np.complex128
    st = np.zeros(1 << n, dtype=dt)
    out = C.c_void_p()
    rc = lib().orc_run(text.encode(), st.ctypes.data, st.nbytes, C.byref(out))
    if rc != 0:
        raise RuntimeError(f"orc_run failed: {rc}")
    res = C.string_at(out.value).decode()
    lib().orc_free(out)
    return st, [l for l in res.splitlines() if l]


def _sfx(a: np.ndarray) -> str:
    return "sp" if a.dtype == np.complex64 else "dp"


def _n(a: np.ndarray) -> int:
    return int(a.size).bit_length() - 1


def _mat(a: np.ndarray, m) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(m, dtype=a.dtype).reshape(-1))


# Kernel-level entry points on numpy states (positions, not program qubits).
def apply_1q(st: np.ndarray, pos: int, m) -> None:
    mm = _mat(st, m)
    getattr(lib(), f"orc_1q_{_sfx(st)}")(st.ctypes.data, _n(st), pos, mm.ctypes.data)


def apply_c1q(st: np.ndarray, c: int, t: int, m) -> None:
    mm = _mat(st, m)
    getattr(lib(), f"orc_c1q_{_sfx(st)}")(st.ctypes.data, _n(st), c, t, mm.ctypes.data)


def apply_swaplike(st: np.ndarray, p1: int, p2: int, m) -> None:
    mm = _mat(st, m)
    getattr(lib(), f"orc_swaplike_{_sfx(st)}")(st.ctypes.data, _n(st), p1, p2, mm.ctypes.data)


def apply_diag(st: np.ndarray, p1: int, p2: int, d) -> None:
    dd = _mat(st, d)
    getattr(lib(), f"orc_diag_{_sfx(st)}")(st.ctypes.data, _n(st), p1, p2, dd.ctypes.data)


def apply_2q(st: np.ndarray, ph: int, pl: int, m) -> None:
    mm = _mat(st, m)
    getattr(lib(), f"orc_2q_{_sfx(st)}")(st.ctypes.data, _n(st), ph, pl, mm.ctypes.data)


def permute(st: np.ndarray, old_map, new_imap) -> None:
    om = np.ascontiguousarray(old_map, dtype=np.uint64)
    ni = np.ascontiguousarray(new_imap, dtype=np.uint64)
    getattr(lib(), f"orc_permute_{_sfx(st)}")(st.ctypes.data, _n(st), om.ctypes.data, ni.ctypes.data)


def prob1(st: np.ndarray, pos: int) -> float:
    return getattr(lib(), f"orc_prob1_{_sfx(st)}")(st.ctypes.data, _n(st), pos)


def norm(st: np.ndarray) -> float:
    return getattr(lib(), f"orc_norm_{_sfx(st)}")(st.ctypes.data, _n(st))


def parity_norm(st: np.ndarray, ymask: int) -> float:
    return getattr(lib(), f"orc_parity_norm_{_sfx(st)}")(st.ctypes.data, _n(st), ymask)


def overlap(a: np.ndarray, b: np.ndarray) -> complex:
    re, im = C.c_double(), C.c_double()
    getattr(lib(), f"orc_overlap_{_sfx(a)}")(a.ctypes.data, b.ctypes.data, _n(a), C.byref(re), C.byref(im))
    return complex(re.value, im.value)


def collapse(st: np.ndarray, pos: int, value: int) -> None:
    getattr(lib(), f"orc_collapse_{_sfx(st)}")(st.ctypes.data, _n(st), pos, int(value))


def qaoa_phase(psi: np.ndarray, diag: np.ndarray, gamma: float) -> None:
    getattr(lib(), f"orc_qaoa_phase_{_sfx(psi)}")(psi.ctypes.data, diag.ctypes.data, _n(psi), gamma)


def qaoa_expect(psi: np.ndarray, diag: np.ndarray) -> float:
    return getattr(lib(), f"orc_qaoa_expect_{_sfx(psi)}")(psi.ctypes.data, diag.ctypes.data, _n(psi))
