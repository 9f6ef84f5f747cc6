"""iqsh_expval_pauli_noalloc (partial sums in a fixed device array, no allocation) against
iqsh_expval_pauli (the library's allocated reduction workspace): the same kernels in the
same order, so the two agree bit for bit.  Two calls in a row, the second on a smaller
register, for every Pauli-string kind (Z only, X only, Y factors), in both precisions;
then repeated calls alternating between the two entry points."""
import ctypes as C

import numpy as np
import pytest

from iqs_amd._lib import IqsError, call
from iqs_amd.engine import StateVector

pytestmark = pytest.mark.gpu


def rand_state(n, dt, seed):
    r = np.random.default_rng(seed)
    v = r.normal(size=1 << n) + 1j * r.normal(size=1 << n)
    v /= np.linalg.norm(v)
    return v.astype(np.complex128 if dt == "c128" else np.complex64)


def both(sv, xm, zm):
    ny = bin(xm & zm).count("1")
    a, an, b, bn = (C.c_double() for _ in range(4))
    call("iqsh_expval_pauli", sv.handle, xm, zm, ny, C.byref(a), C.byref(an))
    call("iqsh_expval_pauli_noalloc", sv.handle, xm, zm, ny, C.byref(b), C.byref(bn))
    return (a.value, an.value), (b.value, bn.value)


def dense_expval(h, xm, zm):
    """<h|P|h> on the host: P|j> = i^ny (-1)^popc(j & z) |j ^ x> for P with X-or-Y
    positions xm and Y-or-Z positions zm."""
    a = h.astype(np.complex128)
    j = np.arange(a.size)
    ny = bin(xm & zm).count("1")
    par = np.array([bin(int(k) & zm).count("1") & 1 for k in j])
    pa = np.zeros_like(a)
    pa[j ^ xm] = (1j ** ny) * np.where(par, -1.0, 1.0) * a
    return float(np.vdot(a, pa).real)


@pytest.mark.parametrize("dt", ["c128", "c64"])
def test_noalloc_matches_allocating_entry_point(dt):
    print(f"no-alloc expectation: {dt}", flush=True)
    for n, seed in ((22, 1), (14, 2)):  # the second call on a smaller register
        h = rand_state(n, dt, seed)
        sv = StateVector(n, dt)
        try:
            sv.upload(h)
            for xm, zm in ((0, 1 | (1 << (n - 1))), (1 << 3, 0), (3, 3), ((1 << 2) | (1 << (n - 1)), 1 << (n - 1))):
                ref, got = both(sv, xm, zm)
                assert got == ref, (n, xm, zm, got, ref)
                if n == 14:
                    tol = 1e-12 if dt == "c128" else 1e-5
                    assert abs(got[0] - dense_expval(h, xm, zm)) < tol
                    assert abs(got[1] - 1.0) < tol
        finally:
            sv.close()


def test_noalloc_repeated_and_checked():
    sv = StateVector(18, "c128")
    try:
        sv.upload(rand_state(18, "c128", 5))
        first = both(sv, 1 << 17, 1)
        for _ in range(5):
            assert both(sv, 1 << 17, 1) == first
        with pytest.raises(IqsError):
            call("iqsh_expval_pauli_noalloc", sv.handle, 1 << 18, 0, 0, None, None)  # mask beyond register
        with pytest.raises(IqsError):
            call("iqsh_expval_pauli_noalloc", sv.handle, 3, 3, 1, None, None)  # ny != popcount(x & z)
    finally:
        sv.close()
