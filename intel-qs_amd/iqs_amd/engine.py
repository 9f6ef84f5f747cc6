"""Python view of one device-resident state (positions-level API over the C ABI).

`StateVector` owns (or wraps) 2^n amplitudes in HBM and exposes the kernels of
libiqs_hip.so one to one.  Program-qubit bookkeeping (permutation maps, named gates) lives
in the iqs::QubitRegister drop-in; this class is the thin layer used by the drop-in's
Python tests, the autograd functions and the benchmarks.
"""
from __future__ import annotations

import ctypes as C
from typing import Iterable, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._lib import C64, C128, call

_DT = {"c64": C64, "c128": C128, np.complex64: C64, np.complex128: C128}


def _cmat(m, size: int) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(m, dtype=np.complex128).reshape(-1))
    if a.size != size:
        raise ValueError(f"expected {size} complex entries, got {a.size}")
    return a


def _edges(edges) -> Tuple[int, object, object, object, tuple]:
    e = list(edges)
    u = (C.c_uint * max(len(e), 1))(*[int(x[0]) for x in e])
    v = (C.c_uint * max(len(e), 1))(*[int(x[1]) for x in e])
    w = (C.c_double * max(len(e), 1))(*[float(x[2]) if len(x) > 2 else 1.0 for x in e])
    return len(e), u, v, w, (u, v, w)


class StateVector:
    def __init__(self, num_qubits: int, dtype="c128", *, _handle=None):
        self.num_qubits = int(num_qubits)
        self.dtype_code = _DT[dtype] if not isinstance(dtype, int) else dtype
        self.np_dtype = np.complex128 if self.dtype_code == C128 else np.complex64
        if _handle is not None:
            self._h = _handle
        else:
            h = C.c_void_p()
            call("iqsh_create", self.num_qubits, self.dtype_code, C.byref(h))
            self._h = h

    @classmethod
    def wrap(cls, num_qubits: int, dtype, device_ptr: int) -> "StateVector":
        h = C.c_void_p()
        code = _DT[dtype] if not isinstance(dtype, int) else dtype
        call("iqsh_wrap", int(num_qubits), code, C.c_void_p(device_ptr), C.byref(h))
        return cls(num_qubits, code, _handle=h)

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            call("iqsh_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def size(self) -> int:
        return 1 << self.num_qubits

    def device_ptr(self) -> int:
        p = C.c_void_p()
        call("iqsh_info", self._h, None, None, C.byref(p))
        return p.value

    # ---- data movement / init ----
    def upload(self, host: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(host, dtype=self.np_dtype)
        call("iqsh_upload", self._h, a.ctypes.data, offset, a.size)

    def download(self, offset: int = 0, count: Optional[int] = None) -> np.ndarray:
        count = self.size - offset if count is None else count
        out = np.empty(count, dtype=self.np_dtype)
        call("iqsh_download", self._h, out.ctypes.data, offset, count)
        return out

    def copy_from(self, other: "StateVector") -> None:
        call("iqsh_copy", self._h, other._h)

    def init_basis(self, index: int) -> None:
        call("iqsh_init_basis", self._h, int(index))

    def init_uniform(self, value: complex) -> None:
        call("iqsh_init_uniform", self._h, float(np.real(value)), float(np.imag(value)))

    # ---- gates (positions) ----
    def apply_1q(self, pos: int, m) -> None:
        a = _cmat(m, 4)
        call("iqsh_apply_1q", self._h, pos, a.ctypes.data)

    def apply_1q_all(self, m) -> None:
        """The same 2x2 matrix on every qubit (a few fused passes; >= 10 qubits)."""
        a = _cmat(m, 4)
        call("iqsh_apply_1q_all", self._h, a.ctypes.data)

    def apply_1q_spec(self, pos: int, spec: int, angle: float = 0.0) -> None:
        call("iqsh_apply_1q_spec", self._h, pos, spec, float(angle))

    def apply_c1q(self, c: int, t: int, m) -> None:
        a = _cmat(m, 4)
        call("iqsh_apply_c1q", self._h, c, t, a.ctypes.data)

    def apply_c1q_spec(self, c: int, t: int, spec: int, angle: float = 0.0) -> None:
        call("iqsh_apply_c1q_spec", self._h, c, t, spec, float(angle))

    def apply_mc1q(self, controls: Iterable[int], t: int, m) -> None:
        mask = 0
        for c in controls:
            mask |= 1 << int(c)
        a = _cmat(m, 4)
        call("iqsh_apply_mc1q", self._h, mask, t, a.ctypes.data)

    def apply_swaplike(self, p1: int, p2: int, m) -> None:
        a = _cmat(m, 4)
        call("iqsh_apply_swaplike", self._h, p1, p2, a.ctypes.data)

    def apply_diag2(self, p1: int, p2: int, d) -> None:
        a = _cmat(d, 4)
        call("iqsh_apply_diag2", self._h, p1, p2, a.ctypes.data)

    def apply_2q(self, ph: int, pl: int, m) -> None:
        a = _cmat(m, 16)
        call("iqsh_apply_2q", self._h, ph, pl, a.ctypes.data)

    def apply_kq(self, positions: Sequence[int], U) -> None:
        k = len(positions)
        a = _cmat(U, 1 << (2 * k))
        p = (C.c_uint * k)(*[int(x) for x in positions])
        call("iqsh_apply_kq", self._h, k, p, a.ctypes.data)

    def scale(self, f: complex) -> None:
        call("iqsh_scale", self._h, float(np.real(f)), float(np.imag(f)))

    def axpy(self, x: "StateVector", f: complex = 1.0) -> None:
        call("iqsh_axpy", self._h, x._h, float(np.real(f)), float(np.imag(f)))

    def collapse(self, pos: int, value: int) -> None:
        call("iqsh_collapse", self._h, pos, int(value))

    def permute(self, src_pos: Sequence[int]) -> None:
        p = (C.c_uint * self.num_qubits)(*[int(x) for x in src_pos])
        call("iqsh_permute", self._h, p)

    # ---- reductions ----
    def prob1(self, pos: int) -> float:
        r = C.c_double()
        call("iqsh_prob1", self._h, pos, C.byref(r))
        return r.value

    def norm2(self) -> float:
        r = C.c_double()
        call("iqsh_norm2", self._h, C.byref(r))
        return r.value

    def expval_pauli(self, xmask: int, zmask: int) -> Tuple[float, float]:
        v, nn = C.c_double(), C.c_double()
        ny = bin(xmask & zmask).count("1")
        call("iqsh_expval_pauli", self._h, xmask, zmask, ny, C.byref(v), C.byref(nn))
        return v.value, nn.value

    def overlap(self, x: "StateVector") -> complex:
        re, im = C.c_double(), C.c_double()
        call("iqsh_overlap", self._h, x._h, C.byref(re), C.byref(im))
        return complex(re.value, im.value)

    def maxabsdiff(self, x: "StateVector", s: complex = 1.0) -> float:
        r = C.c_double()
        call("iqsh_maxabsdiff", self._h, x._h, float(np.real(s)), float(np.imag(s)), C.byref(r))
        return r.value

    def l2diff2(self, x: "StateVector") -> float:
        r = C.c_double()
        call("iqsh_l2diff2", self._h, x._h, C.byref(r))
        return r.value

    # ---- QAOA (edges as (u, v[, w]) in data positions) ----
    def qaoa_phase(self, edges, gamma: float) -> None:
        ne, u, v, w, _keep = _edges(edges)
        call("iqsh_qaoa_phase", self._h, ne, u, v, w, float(gamma))

    def qaoa_expect(self, edges) -> float:
        ne, u, v, w, _keep = _edges(edges)
        r = C.c_double()
        call("iqsh_qaoa_expect", self._h, ne, u, v, w, C.byref(r))
        return r.value

    def maxcut_diag(self, edges) -> float:
        ne, u, v, w, _keep = _edges(edges)
        r = C.c_double()
        call("iqsh_maxcut_diag", self._h, ne, u, v, w, C.byref(r))
        return r.value

    def qaoa_phase_diag(self, diag: "StateVector", gamma: float) -> None:
        call("iqsh_qaoa_phase_diag", self._h, diag._h, float(gamma))

    def qaoa_expect_diag(self, diag: "StateVector", squared: bool = False) -> float:
        r = C.c_double()
        call("iqsh_qaoa_expect_diag", self._h, diag._h, int(squared), C.byref(r))
        return r.value

    def qaoa_cost_apply(self, lam: "StateVector", edges) -> float:
        ne, u, v, w, _keep = _edges(edges)
        r = C.c_double()
        call("iqsh_qaoa_cost_apply", self._h, lam._h, ne, u, v, w, C.byref(r))
        return r.value

    def qaoa_phase_adjoint(self, lam: "StateVector", edges, gamma: float) -> float:
        ne, u, v, w, _keep = _edges(edges)
        r = C.c_double()
        call("iqsh_qaoa_phase_adjoint", self._h, lam._h, ne, u, v, w, float(gamma), C.byref(r))
        return r.value

    def qaoa_layer(self, edges, gamma: float, mixer) -> None:
        """exp(-i gamma C) then the same mixer on every qubit, fused (>= 10 qubits)."""
        ne, u, v, w, _keep = _edges(edges)
        a = _cmat(mixer, 4)
        call("iqsh_qaoa_layer", self._h, ne, u, v, w, float(gamma), a.ctypes.data)

    def qaoa_layer_adjoint(self, lam: "StateVector", edges, gamma: float, mixer_inv, G) -> Tuple[float, float]:
        """Backward step through one QAOA layer: (d/dgamma, sum_q 2 Im <lam|G_q|psi>)."""
        ne, u, v, w, _keep = _edges(edges)
        a, g = _cmat(mixer_inv, 4), _cmat(G, 4)
        dg, dm = C.c_double(), C.c_double()
        call("iqsh_qaoa_layer_adjoint", self._h, lam._h, ne, u, v, w, float(gamma), a.ctypes.data, g.ctypes.data,
             C.byref(dg), C.byref(dm))
        return dg.value, dm.value

    def adjoint_1q_all(self, lam: "StateVector", uinv, G) -> float:
        """Backward step through a layer of the same gate on every qubit: returns
        sum_q 2 Im <lam| G_q |psi>, then applies uinv to every qubit of both states."""
        a, g = _cmat(uinv, 4), _cmat(G, 4)
        r = C.c_double()
        call("iqsh_adjoint_1q_all", self._h, lam._h, a.ctypes.data, g.ctypes.data, C.byref(r))
        return r.value

    def adjoint_1q(self, lam: "StateVector", pos: int, uinv, G=None) -> Optional[float]:
        a = _cmat(uinv, 4)
        if G is None:
            call("iqsh_adjoint_1q", self._h, lam._h, pos, a.ctypes.data, None, None)
            return None
        g = _cmat(G, 4)
        r = C.c_double()
        call("iqsh_adjoint_1q", self._h, lam._h, pos, a.ctypes.data, g.ctypes.data, C.byref(r))
        return r.value


def sync() -> None:
    call("iqsh_sync")
