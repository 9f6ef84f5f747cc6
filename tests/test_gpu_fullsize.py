"""Parity at the benchmark sizes: 30-qubit complex128 (16 GiB), 32-qubit complex64 (32 GiB)
and, for fusion, 32-qubit complex128 (64 GiB) registers.

At these sizes every launch takes the two-dimensional grid (more than 2^16 workgroups, so
blockIdx.y > 0) with the XCD-contiguous block mapping, and the 5-qubit blocks, permutations
and QAOA layers take their large-register paths.  A full state cannot be checked against a
CPU run in seconds, so the checks are size-independent:

- closed forms: a circuit that entangles only small clusters of positions (controls and
  targets at positions up to n - 1 and below the cache-line width) keeps the state a
  product of cluster states; tests/_factored.py advances each cluster with the oracle, so
  sampled amplitudes, GetProbability of every position, Pauli expectations and the norm
  have exact expected values (relative tolerance 1e-12 complex128, 2e-5 complex64 on
  amplitudes scaled by 2^(n/2); 1e-12 / 1e-5 absolute on reductions);
- the exact inverse sequence returns the state to |0...0> (max-abs 1e-12 / 1e-5);
- moves (X, CNOT, SWAP, Toffoli) and PermuteQubits (reversal, rotation, random) are
  bit-exact: sampled amplitudes equal the source amplitudes under the index map, and the
  inverse moves restore the original register exactly (MaxAbsDiff == 0);
- fused 5-qubit blocks against the unfused FP64 run of the same 200 Haar-random gates;
- depth-1 QAOA MaxCut on the 28-vertex 3-regular graph: cost and adjoint gradient against
  the closed form in tests/_qaoa_p1.py.
References: src/highperfkernels.cpp:287-380 (pair loops), src/qureg_apply1qubitgate.cpp:
184-201, src/qureg_permute.cpp:57-105, src/qureg_fusion.cpp:56-94,
src/qaoa_features.cpp:248-295.
"""
import ctypes as C
import math

import numpy as np
import pytest

import _qaoa_p1
from _factored import Factored
from iqs_amd import gates as G
from iqs_amd._lib import IqsError, call
from iqs_amd.circuits import SplitMix64, haar_unitary, random_1q2q_circuit, random_regular_graph
from iqs_amd.engine import StateVector
from iqs_amd.qaoa import MaxCutQaoa

pytestmark = pytest.mark.gpu

SIZES = [("c128", 30), ("c64", 32)]
REL = {"c128": 1e-12, "c64": 2e-5}   # amplitude error x 2^(n/2)
ABS = {"c128": 1e-12, "c64": 1e-5}   # reductions, |0> round trip


def sample_indices(n, seed):
    r = np.random.default_rng(seed)
    singles = r.integers(0, 1 << n, size=2048, dtype=np.uint64)
    runs = [np.arange(s, s + 64, dtype=np.uint64) for s in r.integers(0, (1 << n) - 64, size=16, dtype=np.uint64)]
    return np.concatenate([singles] + runs)


def sample(sv, idx):
    out = np.empty(len(idx), dtype=sv.np_dtype)
    for k, i in enumerate(idx):
        out[k] = sv.download(int(i), 1)[0]
    return out


def dag(m):
    return np.conj(np.asarray(m, dtype=np.complex128)).T


def cluster_ops(n, seed):
    """Every gate kind, at the top positions and below the line/wavefront widths, grouped so
    that no cluster exceeds 7 positions."""
    r = SplitMix64(seed)
    H = n - 1
    u2 = lambda: haar_unitary(2, r)  # noqa: E731
    u4 = lambda: haar_unitary(4, r)  # noqa: E731
    ops = [("1q", p, u2()) for p in range(n)]
    ops += [
        ("c1q", H, 0, u2()), ("c1q", 0, H, u2()),                 # {0, H}
        ("c1q", 1, H - 1, u2()), ("c1q", H - 1, 2, u2()),         # control in a line; lane target
        ("swap", 3, H - 2, u2()), ("swap", H - 2, H - 3, u2()),
        ("diag2", H - 4, 4, np.exp(1j * np.array([0.3, -1.1, 0.7, 2.0]))), ("2q", H - 4, H - 5, u4()),
        ("2q", H - 6, 5, u4()), ("2q", 6, 7, u4()),
        ("mc1q", (H - 7, H - 8), 8, u2()),
        ("1q", H - 9, G.rz(0.7)), ("1q", H - 10, G.T), ("1q", H - 11, G.X), ("1q", H - 12, G.Z), ("1q", 9, G.SQRT_Z),
        ("c1q", H - 13, H - 14, G.Z), ("c1q", 9, H - 15, G.cphase(0.3)),
        ("kq", (H, H - 2, 0, H - 3, 3), haar_unitary(32, r)),      # five positions, unsorted
        ("kq", (10, 11, 12, 13, H - 1), haar_unitary(32, r)),
        ("kq", (H - 5, 4, H - 4), haar_unitary(8, r)),            # padded to five
        ("all", u2()),
    ]
    return ops


def apply(target, op, inverse=False):
    kind = op[0]
    if kind in ("1q", "all"):
        m = dag(op[-1]) if inverse else op[-1]
        if kind == "all":
            if isinstance(target, Factored):
                for p in range(target.n):
                    target.apply_1q(p, m)
            else:
                target.apply_1q_all(m)
        else:
            target.apply_1q(op[1], m)
    elif kind == "c1q":
        target.apply_c1q(op[1], op[2], dag(op[3]) if inverse else op[3])
    elif kind == "mc1q":
        target.apply_mc1q(op[1], op[2], dag(op[3]) if inverse else op[3])
    elif kind == "swap":
        target.apply_swaplike(op[1], op[2], dag(op[3]) if inverse else op[3])
    elif kind == "diag2":
        target.apply_diag2(op[1], op[2], np.conj(op[3]) if inverse else op[3])
    elif kind == "2q":
        target.apply_2q(op[1], op[2], dag(op[3]) if inverse else op[3])
    elif kind == "kq":
        target.apply_kq(list(op[1]), dag(op[2]) if inverse else op[2])


@pytest.mark.parametrize("dt,n", SIZES)
def test_cluster_circuit_closed_form_and_inverse(dt, n):
    print(f"fullsize closed form: {dt} n={n}", flush=True)
    ops = cluster_ops(n, 100 + n)
    sv = StateVector(n, dt)
    ref = None
    try:
        sv.init_basis(0)
        f = Factored(n)
        for op in ops:
            apply(sv, op)
            apply(f, op)
        assert max(f.cluster_sizes()) <= 7
        idx = sample_indices(n, n)
        got = sample(sv, idx)
        want = f.amplitudes(idx)
        scale = math.sqrt(float(1 << n))
        assert np.abs(got - want).max() * scale < REL[dt]
        assert abs(sv.norm2() - f.norm2()) < ABS[dt]
        for p in range(n):
            assert abs(sv.prob1(p) - f.prob1(p)) < ABS[dt], p
        H = n - 1
        for xm, zm in [(1 << H, 0), ((1 << H) | (1 << (H - 1)), 1 << (H - 1)), (0, 1 | (1 << (H - 4))),
                       ((1 << 5) | (1 << (H - 6)), 1 << (H - 6)), (1 << 1, (1 << 2) | (1 << (H - 9))),
                       ((1 << 10) | (1 << 11), (1 << 12) | (1 << 11))]:
            v, nn = sv.expval_pauli(xm, zm)
            assert abs(v - f.expval_pauli(xm, zm)) < ABS[dt], (xm, zm)
        for op in reversed(ops):
            apply(sv, op, inverse=True)
        ref = StateVector(n, dt)
        ref.init_basis(0)
        assert sv.maxabsdiff(ref) < ABS[dt]
    finally:
        sv.close()
        if ref is not None:
            ref.close()


def _bit_maps(n):
    H = n - 1

    def x(p):
        return lambda i: i ^ (1 << p)

    def cx(c, t):
        return lambda i: i ^ (((i >> c) & 1) << t)

    def sw(a, b):
        return lambda i: i ^ ((((i >> a) ^ (i >> b)) & 1) * ((1 << a) | (1 << b)))

    def ccx(c1, c2, t):
        return lambda i: i ^ ((((i >> c1) & (i >> c2)) & 1) << t)

    moves = [
        ("1q", (0,), x(0)), ("1q", (H,), x(H)), ("1q", (3,), x(3)),
        ("c1q", (H, 0), cx(H, 0)), ("c1q", (0, H), cx(0, H)), ("c1q", (2, H - 1), cx(2, H - 1)),
        ("c1q", (H - 1, H - 2), cx(H - 1, H - 2)), ("c1q", (H - 3, 1), cx(H - 3, 1)),
        ("swap", (0, H), sw(0, H)), ("swap", (4, H - 3), sw(4, H - 3)), ("swap", (H - 1, H), sw(H - 1, H)),
        ("swap", (1, 2), sw(1, 2)),
        ("mc1q", ((H, 1), 5), ccx(H, 1, 5)), ("mc1q", ((0, 2), H - 4), ccx(0, 2, H - 4)),
    ]
    return moves


def _apply_move(sv, mv):
    kind, a, _ = mv
    if kind == "1q":
        sv.apply_1q(a[0], G.X)
    elif kind == "c1q":
        sv.apply_c1q(a[0], a[1], G.X)
    elif kind == "swap":
        sv.apply_swaplike(a[0], a[1], G.X)
    else:
        sv.apply_mc1q(a[0], a[1], G.X)


@pytest.mark.parametrize("dt,n", SIZES)
def test_moves_and_permutations_bit_exact(dt, n):
    print(f"fullsize moves: {dt} n={n}", flush=True)
    r = SplitMix64(7 * n)
    sv, ref = StateVector(n, dt), StateVector(n, dt)
    try:
        sv.init_basis(0)
        for p in range(n):
            sv.apply_1q(p, haar_unitary(2, r))
        sv.apply_c1q(n - 1, 0, haar_unitary(2, r))
        sv.apply_2q(n - 2, 1, haar_unitary(4, r))
        ref.copy_from(sv)
        idx = sample_indices(n, 3 * n)
        moves = _bit_maps(n)
        for mv in moves:
            _apply_move(sv, mv)
        src = idx.copy()
        for mv in reversed(moves):  # new[j] = old[g^-1(j)], every move an involution
            src = np.array([mv[2](int(i)) for i in src], dtype=np.uint64)
        assert np.array_equal(sample(sv, idx), sample(ref, src))
        for mv in reversed(moves):
            _apply_move(sv, mv)
        assert sv.maxabsdiff(ref) == 0.0
        rnd = list(np.random.default_rng(n).permutation(n))
        for perm in (list(reversed(range(n))), [(b + 1) % n for b in range(n)], rnd):
            print(f"  permute {perm}", flush=True)
            sv.permute(perm)
            src = np.array([sum(((int(j) >> b) & 1) << int(perm[b]) for b in range(n)) for j in idx], dtype=np.uint64)
            assert np.array_equal(sample(sv, idx), sample(ref, src))
            inv = [0] * n
            for b, s in enumerate(perm):
                inv[s] = b
            sv.permute(inv)
            assert sv.maxabsdiff(ref) == 0.0
    finally:
        sv.close()
        ref.close()


def test_fused_blocks_32q_against_unfused_fp64():
    """Config 3 as a test: 200 Haar-random 1q/2q gates on 32 qubits fused into <= 5-qubit
    blocks (complex128 and complex64) against the unfused complex128 run."""
    import intelqs_py as iqs
    from iqs_amd import runner

    n = 32
    print("fullsize fusion: 32 qubits", flush=True)

    def h(reg):
        return C.c_void_p(reg.DeviceHandle())

    c = random_1q2q_circuit(n, 200, 20240603)
    ref = iqs.QubitRegister(n, "base", 0, 0)
    runner.run(c, ref)
    scale = math.sqrt(float(1 << n))
    for prec, cls, rel, inf_tol in (("dp", iqs.QubitRegister, 1e-12, 1e-12), ("sp", iqs.QubitRegisterSP, 1e-4, 1e-9)):
        c.precision = prec
        reg = cls(n, "base", 0, 0)
        reg.SetFusionBlockQubits(5)
        reg.TurnOnFusion()
        runner.run(c, reg)
        gates, blocks = reg.FusionStats()
        assert gates == 200 and blocks <= 26  # dependency-graph planner (59 when closed at the first misfit)
        a = h(reg)
        tmp = None
        if prec == "sp":
            tmp = C.c_void_p()
            call("iqsh_create", n, 1, C.byref(tmp))
            call("iqsh_convert", tmp, a)
            a = tmp
        dre, dim, n2a, n2b, mx = (C.c_double() for _ in range(5))
        call("iqsh_overlap", a, h(ref), C.byref(dre), C.byref(dim))
        call("iqsh_norm2", a, C.byref(n2a))
        call("iqsh_norm2", h(ref), C.byref(n2b))
        call("iqsh_maxabsdiff", a, h(ref), 1.0, 0.0, C.byref(mx))
        if tmp is not None:
            call("iqsh_destroy", tmp)
        del reg
        inf = 1.0 - (dre.value ** 2 + dim.value ** 2) / (n2a.value * n2b.value)
        print(f"  {prec}: blocks {blocks}, 1-F {inf:.3e}, max-abs x 2^16 {mx.value * scale:.3e}", flush=True)
        assert abs(inf) < inf_tol
        assert mx.value * scale < rel
    c.precision = "dp"


@pytest.mark.parametrize("dt", ["c128", "c64"])
def test_qaoa_depth1_28q_closed_form(dt):
    """Cost and adjoint gradient of depth-1 QAOA on the 28-vertex random 3-regular graph
    (seed 20240604) against the closed form, at several angles."""
    n = 28
    print(f"fullsize qaoa p=1: {dt}", flush=True)
    edges = random_regular_graph(n, 3, 20240604)
    prob = MaxCutQaoa(n, edges, dt)
    tol_c, tol_g = (1e-10, 1e-9) if dt == "c128" else (2e-4, 2e-4)
    for g, b in [(0.37, 0.21), (1.3, 2.5), (2.9, 0.8)]:
        want = _qaoa_p1.cost(n, edges, g, b)
        wg, wb = _qaoa_p1.grad(n, edges, g, b)
        cost, dg, db = prob.adjoint([g], [b])
        assert abs(prob.cost([g], [b]) - want) < tol_c
        assert abs(cost - want) < tol_c
        assert abs(dg[0] - wg) < tol_g and abs(db[0] - wb) < tol_g, (dg, wg, db, wb)


def test_oversized_register_fails_cleanly():
    """A 36-qubit complex128 register (1 TiB) does not fit: iqsh_create reports
    IQSH_ERR_ALLOC with the hipMalloc error, and the library keeps working."""
    with pytest.raises(IqsError, match="error 2: hipMalloc"):
        StateVector(36, "c128")
    sv = StateVector(12, "c64")
    sv.init_basis(5)
    assert sv.prob1(0) == 1.0 and sv.prob1(1) == 0.0
    sv.close()
