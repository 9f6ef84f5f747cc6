"""Kernel-level parity on the GPU: every C-ABI entry point against the CPU oracle on the
same seeded input.  Tolerances: 1e-12 max-abs (complex128), 1e-5 (complex64); moves and
permutations bit-exact; reductions against FP64 sums."""
import math

import numpy as np
import pytest

import oracle
from iqs_amd.circuits import SplitMix64, haar_unitary, random_su2
from iqs_amd.engine import StateVector

pytestmark = pytest.mark.gpu

DT = {"c128": np.complex128, "c64": np.complex64}
TOL = {"c128": 1e-12, "c64": 1e-5}


def rand_state(n, dt, seed):
    r = np.random.default_rng(seed)
    v = r.normal(size=1 << n) + 1j * r.normal(size=1 << n)
    return (v / np.linalg.norm(v)).astype(DT[dt])


def on_gpu(host, dt):
    sv = StateVector(int(host.size).bit_length() - 1, dt)
    sv.upload(host)
    return sv


@pytest.fixture(params=["c128", "c64"])
def dt(request):
    return request.param


def test_upload_download_roundtrip(dt):
    h = rand_state(12, dt, 1)
    sv = on_gpu(h, dt)
    assert np.array_equal(sv.download(), h)
    assert np.array_equal(sv.download(100, 7), h[100:107])


@pytest.mark.parametrize("pos", range(12))
def test_1q_dense_every_position(dt, pos):
    n = 12
    m = haar_unitary(2, SplitMix64(pos + 11))
    h = rand_state(n, dt, pos)
    sv = on_gpu(h, dt)
    sv.apply_1q(pos, m)
    oracle.apply_1q(h, pos, m)
    assert np.abs(sv.download() - h).max() < TOL[dt]


@pytest.mark.parametrize("kind", ["X", "Y", "Z", "S", "T", "RZ", "H", "I"])
def test_1q_structured(dt, kind):
    n, pos = 11, 5
    mats = {
        "X": [[0, 1], [1, 0]], "Y": [[0, -1j], [1j, 0]], "Z": [[1, 0], [0, -1]],
        "S": [[1, 0], [0, 1j]], "T": [[1, 0], [0, np.exp(1j * math.pi / 4)]],
        "RZ": [[np.exp(-0.3j), 0], [0, np.exp(0.3j)]],
        "H": np.array([[1, 1], [1, -1]]) / math.sqrt(2), "I": [[1, 0], [0, 1]],
    }
    m = np.array(mats[kind], dtype=complex)
    h = rand_state(n, dt, 3)
    sv = on_gpu(h, dt)
    sv.apply_1q(pos, m)
    oracle.apply_1q(h, pos, m)
    got = sv.download()
    if kind in ("X", "I"):
        assert np.array_equal(got, h)  # pure move / no-op
    else:
        assert np.abs(got - h).max() < TOL[dt]


@pytest.mark.parametrize("c,t", [(0, 1), (1, 0), (0, 10), (10, 0), (3, 7), (7, 3), (9, 10)])
def test_c1q(dt, c, t):
    n = 11
    m = haar_unitary(2, SplitMix64(c * 13 + t))
    h = rand_state(n, dt, c + t)
    sv = on_gpu(h, dt)
    sv.apply_c1q(c, t, m)
    oracle.apply_c1q(h, c, t, m)
    assert np.abs(sv.download() - h).max() < TOL[dt]


def test_cnot_and_cz_exact(dt):
    n = 10
    h = rand_state(n, dt, 9)
    sv = on_gpu(h, dt)
    X = np.array([[0, 1], [1, 0]], dtype=complex)
    Z = np.array([[1, 0], [0, -1]], dtype=complex)
    for (c, t) in [(0, 1), (5, 2), (9, 0)]:
        sv.apply_c1q(c, t, X)
        oracle.apply_c1q(h, c, t, X)
    assert np.array_equal(sv.download(), h)
    sv.apply_c1q(3, 8, Z)
    oracle.apply_c1q(h, 3, 8, Z)
    assert np.array_equal(sv.download(), h)


def test_multi_controlled_toffoli(dt):
    n = 9
    h = rand_state(n, dt, 21)
    sv = on_gpu(h, dt)
    X = np.array([[0, 1], [1, 0]], dtype=complex)
    sv.apply_mc1q([2, 7], 4, X)
    want = h.copy()
    for i in range(1 << n):
        if (i >> 2) & 1 and (i >> 7) & 1 and not (i >> 4) & 1:
            j = i | (1 << 4)
            want[i], want[j] = h[j], h[i]
    assert np.array_equal(sv.download(), want)


@pytest.mark.parametrize("p1,p2", [(0, 1), (1, 0), (0, 10), (4, 6), (10, 3)])
def test_swaplike(dt, p1, p2):
    n = 11
    h = rand_state(n, dt, 5)
    sv = on_gpu(h, dt)
    th = 0.37 + p1
    m = np.array([[math.cos(th), 1j * math.sin(th)], [1j * math.sin(th), math.cos(th)]])
    sv.apply_swaplike(p1, p2, m)
    oracle.apply_swaplike(h, min(p1, p2), max(p1, p2), m)
    assert np.abs(sv.download() - h).max() < TOL[dt]
    sv.apply_swaplike(p1, p2, [[0, 1], [1, 0]])
    oracle.apply_swaplike(h, min(p1, p2), max(p1, p2), np.array([[0, 1], [1, 0]]))
    assert np.abs(sv.download() - h).max() < TOL[dt]


@pytest.mark.parametrize("p1,p2", [(0, 1), (1, 0), (2, 9), (9, 2)])
@pytest.mark.parametrize("kind", ["general", "cz", "cphase01"])
def test_diag2(dt, p1, p2, kind):
    n = 10
    d = {"general": np.exp(1j * np.array([0.2, -1.1, 0.7, 2.5])),
         "cz": np.array([1, 1, 1, -1], dtype=complex),
         "cphase01": np.array([1, np.exp(0.4j), 1, 1])}[kind]
    h = rand_state(n, dt, 8)
    sv = on_gpu(h, dt)
    sv.apply_diag2(p1, p2, d)
    oracle.apply_diag(h, p1, p2, d)
    assert np.abs(sv.download() - h).max() < TOL[dt]


@pytest.mark.parametrize("ph,pl", [(1, 0), (0, 1), (9, 0), (0, 9), (5, 6), (8, 3)])
def test_2q_dense(dt, ph, pl):
    n = 10
    m = haar_unitary(4, SplitMix64(ph * 17 + pl))
    h = rand_state(n, dt, ph + 2 * pl)
    sv = on_gpu(h, dt)
    sv.apply_2q(ph, pl, m)
    oracle.apply_2q(h, ph, pl, m)
    assert np.abs(sv.download() - h).max() < TOL[dt]


@pytest.mark.parametrize("positions", [[0], [3], [0, 1], [7, 2], [0, 1, 2], [9, 4, 0], [1, 3, 5, 7],
                                       [0, 1, 2, 3, 4], [9, 2, 6, 0, 4]])
def test_kq_block_matches_kronecker(dt, positions):
    n = 10
    k = len(positions)
    U = haar_unitary(1 << k, SplitMix64(sum(positions) + 100 * k))
    h = rand_state(n, dt, k)
    sv = on_gpu(h, dt)
    sv.apply_kq(positions, U)
    # reference: explicit gather/scatter with row bit j <-> positions[j]
    want = h.astype(np.complex128).copy()
    mask = sum(1 << p for p in positions)
    for base in range(1 << n):
        if base & mask:
            continue
        idx = [base | sum(((r >> j) & 1) << positions[j] for j in range(k)) for r in range(1 << k)]
        want[idx] = U @ h.astype(np.complex128)[idx]
    assert np.abs(sv.download() - want).max() < (1e-12 if dt == "c128" else 2e-5)


def _perm_cases():
    """(n, src_pos): random permutations below and above 16 qubits (from 16 up, a
    permutation that moves low positions up runs as a tiled transpose through LDS), the
    reversal, a rotation, and an exchange of high positions only (gather kernel)."""
    cases = [(11, 1), (11, 2), (11, 3), (17, 4), (20, 5)]
    out = [(n, [int(x) for x in np.random.default_rng(s).permutation(n)]) for n, s in cases]
    out.append((18, list(range(17, -1, -1))))
    out.append((18, [(b + 1) % 18 for b in range(18)]))
    out.append((18, [12 if b == 16 else 16 if b == 12 else b for b in range(18)]))
    return out


@pytest.mark.parametrize("case", _perm_cases(), ids=lambda c: f"n{c[0]}-{''.join(chr(97 + x) for x in c[1][:6])}")
def test_permute_bit_exact(dt, case):
    n, src = case
    perm_seed = n
    h = rand_state(n, dt, perm_seed)
    sv = on_gpu(h, dt)
    sv.permute(src)
    j = np.arange(1 << n)
    srcidx = np.zeros_like(j)
    for b in range(n):
        srcidx |= ((j >> b) & 1) << src[b]
    assert np.array_equal(sv.download(), h[srcidx])


def test_reductions(dt):
    n = 12
    h = rand_state(n, dt, 31) * 1.7
    sv = on_gpu(h, dt)
    a = h.astype(np.complex128)
    p = np.abs(a) ** 2
    idx = np.arange(1 << n)
    for pos in (0, 5, 11):
        assert abs(sv.prob1(pos) - p[(idx >> pos) & 1 == 1].sum()) < 1e-12
    assert abs(sv.norm2() - p.sum()) < 1e-12
    g = rand_state(n, dt, 32)
    sg = on_gpu(g, dt)
    assert abs(sv.overlap(sg) - np.vdot(g.astype(np.complex128), a)) < 1e-12
    assert abs(sv.l2diff2(sg) - (np.abs(a - g) ** 2).sum()) < 1e-12
    gg = g.astype(np.complex128)
    assert abs(sv.maxabsdiff(sg, 0.5 - 0.25j) - np.abs(a - (0.5 - 0.25j) * gg).max()) < 1e-12


def pauli_dense(n, paulis):
    """Explicit operator of a Pauli string {position: 'X'|'Y'|'Z'}."""
    P = {"I": np.eye(2), "X": np.array([[0, 1], [1, 0]]), "Y": np.array([[0, -1j], [1j, 0]]),
         "Z": np.diag([1, -1])}
    full = np.array([[1.0 + 0j]])
    for q in reversed(range(n)):
        full = np.kron(full, P[paulis.get(q, "I")])
    return full


@pytest.mark.parametrize("paulis", [{0: "Z"}, {3: "X"}, {2: "Y"}, {0: "X", 1: "X"}, {1: "Y", 4: "Y"},
                                    {0: "X", 2: "Y", 5: "Z"}, {5: "Y", 3: "Y", 1: "Y"}, {4: "Z", 1: "Z"}])
def test_pauli_expectation_one_pass(dt, paulis):
    n = 6
    h = rand_state(n, dt, 77) * 1.1
    sv = on_gpu(h, dt)
    x = sum(1 << q for q, s in paulis.items() if s in "XY")
    z = sum(1 << q for q, s in paulis.items() if s in "YZ")
    val, nn = sv.expval_pauli(x, z)
    a = h.astype(np.complex128)
    want = np.vdot(a, pauli_dense(n, paulis) @ a).real
    assert abs(val - want) < 1e-12
    assert abs(nn - np.vdot(a, a).real) < 1e-12
    assert np.array_equal(sv.download(), h)  # read-only


def pauli_apply(a, n, paulis):
    """P a for a Pauli string, one 2x2 factor per position along its axis of a."""
    P = {"X": np.array([[0, 1], [1, 0]]), "Y": np.array([[0, -1j], [1j, 0]]), "Z": np.diag([1, -1])}
    out = a.copy()
    for q, s in paulis.items():
        v = out.reshape(1 << (n - 1 - q), 2, 1 << q)
        out = np.einsum("ij,ajb->aib", P[s], v).reshape(-1)
    return out


@pytest.mark.parametrize("n,paulis", [
    (4, {0: "X"}), (4, {1: "Y", 3: "X"}), (4, {0: "Y", 2: "Z", 3: "Y"}),
    (13, {0: "X", 1: "Y", 12: "Z"}), (13, {5: "Y", 2: "X", 7: "Z"}), (13, {6: "X"}),
    (13, {0: "X", 12: "Y", 3: "Z"}), (13, {11: "Y", 6: "Y", 9: "Z"})])
def test_pauli_expectation_lane_and_pair_forms(dt, n, paulis):
    """X/Y positions all below 6 (partner by lane shuffle; below 64 items too) and above."""
    assert np.allclose(pauli_apply(np.arange(64.0) + 1j, 6, {0: "X", 2: "Y", 5: "Z"}),
                       pauli_dense(6, {0: "X", 2: "Y", 5: "Z"}) @ (np.arange(64.0) + 1j))
    h = rand_state(n, dt, 78 + n) * 1.1
    sv = on_gpu(h, dt)
    x = sum(1 << q for q, s in paulis.items() if s in "XY")
    z = sum(1 << q for q, s in paulis.items() if s in "YZ")
    val, nn = sv.expval_pauli(x, z)
    a = h.astype(np.complex128)
    want = np.vdot(a, pauli_apply(a, n, paulis)).real
    assert abs(val - want) < 1e-12
    assert abs(nn - np.vdot(a, a).real) < 1e-12


@pytest.mark.parametrize("n", [9, 20])
def test_reductions_repeat_bit_identically(dt, n):
    """The summation order depends on the size only: repeated reductions, with other
    reductions (other grids) between them, return the same bits.  n = 9 leaves every thread
    in the tail of its first chunk; n = 20 runs whole chunks on many workgroups."""
    h = rand_state(n, dt, 90 + n)
    g = rand_state(n, dt, 91 + n)
    sv, sg = on_gpu(h, dt), on_gpu(g, dt)

    def everything():
        return ([sv.norm2()] + [sv.prob1(p) for p in (0, 3, n - 1)]
                + [v for xz in ((0, 5), (3, 6), (1 << (n - 1), 1), ((1 << 7) | 2, 4)) for v in sv.expval_pauli(*xz)]
                + [sv.overlap(sg).real, sv.overlap(sg).imag, sv.maxabsdiff(sg, 1.0)])

    first = everything()
    sg.norm2()
    second = everything()
    assert [float(x).hex() for x in first] == [float(x).hex() for x in second]
    a = h.astype(np.complex128)
    assert abs(first[0] - np.vdot(a, a).real) < 1e-12


def test_scale_axpy_collapse_fill(dt):
    n = 10
    h = rand_state(n, dt, 4)
    g = rand_state(n, dt, 5)
    sv, sg = on_gpu(h, dt), on_gpu(g, dt)
    sv.scale(0.3 - 0.2j)
    sv.axpy(sg, 1.5j)
    want = (0.3 - 0.2j) * h.astype(complex) + 1.5j * g.astype(complex)
    assert np.abs(sv.download() - want).max() < 10 * TOL[dt]
    sv.collapse(3, 1)
    want[((np.arange(1 << n) >> 3) & 1) == 0] = 0
    assert np.abs(sv.download() - want).max() < 10 * TOL[dt]
    sv.init_uniform(0.25)
    assert np.all(sv.download() == 0.25)
    sv.init_basis(77)
    d = sv.download()
    assert d[77] == 1 and np.count_nonzero(d) == 1


def edges_3reg(nv, seed):
    from iqs_amd.circuits import random_regular_graph
    return random_regular_graph(nv, 3, seed)


@pytest.mark.parametrize("weighted", [False, True])
def test_qaoa_phase_and_expectation(dt, weighted):
    n = 10
    edges = edges_3reg(n, 7)
    if weighted:
        edges = [(u, v, 0.5 + 0.1 * k) for k, (u, v, _) in enumerate(edges)]
    h = rand_state(n, dt, 12)
    sv = on_gpu(h, dt)
    idx = np.arange(1 << n)
    cut = np.zeros(1 << n)
    for (u, v, w) in edges:
        cut += w * (((idx >> u) ^ (idx >> v)) & 1)
    gamma = 0.731
    sv.qaoa_phase(edges, gamma)
    want = h.astype(np.complex128) * np.exp(-1j * gamma * cut)
    got = sv.download().astype(np.complex128)
    assert np.abs(got - want).max() < TOL[dt]
    assert abs(sv.qaoa_expect(edges) - (cut * np.abs(got) ** 2).sum()) < 1e-12
    diag = StateVector(n, dt)
    mx = diag.maxcut_diag(edges)
    assert abs(mx - cut.max()) < 1e-12
    assert np.abs(diag.download().real - cut).max() < 1e-5


def test_qaoa_diag_register_matches_oracle(dt):
    n = 8
    edges = edges_3reg(n, 3)
    h = rand_state(n, dt, 2)
    sv = on_gpu(h, dt)
    diag = StateVector(n, dt)
    diag.maxcut_diag(edges)
    dh = diag.download()
    sv.qaoa_phase_diag(diag, 0.4)
    oracle.qaoa_phase(h, dh, 0.4)
    assert np.abs(sv.download() - h).max() < TOL[dt]
    assert abs(sv.qaoa_expect_diag(diag) - oracle.qaoa_expect(h, dh)) < (1e-12 if dt == "c128" else 1e-5)


def test_errors_are_reported_not_fatal():
    from iqs_amd._lib import IqsError
    sv = StateVector(4, "c128")
    with pytest.raises(IqsError, match="out of range"):
        sv.apply_1q(4, np.eye(2))
    with pytest.raises(IqsError, match="coincide"):
        sv.apply_c1q(1, 1, np.eye(2))
    with pytest.raises(IqsError):
        StateVector(41, "c128")
    with pytest.raises(IqsError, match="not a permutation"):
        sv.permute([0, 0, 1, 2])
    other = StateVector(5, "c128")
    with pytest.raises(IqsError, match="differ"):
        sv.overlap(other)


# ---- large enough registers for the tiled (XCD / split-block) and lane-shuffle paths ----
NT = 17


@pytest.fixture(scope="module")
def big_states():
    return {dt: rand_state(NT, dt, 1234) for dt in ("c128", "c64")}


@pytest.mark.parametrize("pos", [0, 1, 3, 5, 6, 7, 10, 11, 12, 13, 16])
def test_tiled_1q_dense(dt, big_states, pos):
    h = big_states[dt].copy()
    m = haar_unitary(2, SplitMix64(900 + pos))
    sv = on_gpu(h, dt)
    sv.apply_1q(pos, m)
    oracle.apply_1q(h, pos, m)
    assert np.abs(sv.download() - h).max() < TOL[dt]


@pytest.mark.parametrize("c,t", [(0, 1), (1, 0), (16, 0), (0, 16), (5, 12), (12, 5), (15, 16)])
def test_tiled_controlled_and_cnot(dt, big_states, c, t):
    h = big_states[dt].copy()
    m = haar_unitary(2, SplitMix64(c * 31 + t))
    sv = on_gpu(h, dt)
    sv.apply_c1q(c, t, m)
    oracle.apply_c1q(h, c, t, m)
    assert np.abs(sv.download() - h).max() < TOL[dt]
    X = np.array([[0, 1], [1, 0]], dtype=complex)
    sv.apply_c1q(t, c, X)
    oracle.apply_c1q(h, t, c, X)
    got = sv.download()
    assert np.abs(got - h).max() < TOL[dt]


@pytest.mark.parametrize("p1,p2", [(0, 1), (0, 16), (7, 9), (15, 2)])
def test_tiled_swap_diag_2q(dt, big_states, p1, p2):
    h = big_states[dt].copy()
    sv = on_gpu(h, dt)
    sv.apply_swaplike(p1, p2, [[0, 1], [1, 0]])
    oracle.apply_swaplike(h, min(p1, p2), max(p1, p2), np.array([[0, 1], [1, 0]]))
    assert np.array_equal(sv.download(), h)
    d = np.exp(1j * np.array([0.1, 0.9, -0.4, 1.7]))
    sv.apply_diag2(p1, p2, d)
    oracle.apply_diag(h, p1, p2, d)
    sv.apply_1q(p2, np.diag([np.exp(-0.2j), np.exp(0.6j)]))
    oracle.apply_1q(h, p2, np.diag([np.exp(-0.2j), np.exp(0.6j)]))
    sv.apply_c1q(p1, p2, np.diag([1, np.exp(0.3j)]))
    oracle.apply_c1q(h, p1, p2, np.diag([1, np.exp(0.3j)]))
    m4 = haar_unitary(4, SplitMix64(p1 + 50 * p2))
    sv.apply_2q(p1, p2, m4)
    oracle.apply_2q(h, p1, p2, m4)
    assert np.abs(sv.download() - h).max() < 4 * TOL[dt]


@pytest.mark.parametrize("p,q", [(0, 9), (1, 0), (2, 16), (3, 1), (5, 2), (12, 3)])
def test_tiled_unit_diagonals_at_low_positions(dt, big_states, p, q):
    """Diagonals with entries exactly 1 (Z/S/T, CPhase, CZ): a position inside a cache line
    goes through the full elementwise pass, elsewhere only the changed amplitudes move."""
    h = big_states[dt].copy()
    sv = on_gpu(h, dt)
    f = np.exp(0.7j)
    for m in (np.diag([1, f]), np.diag([f, 1]), np.diag([1, -1])):
        sv.apply_1q(p, m)
        oracle.apply_1q(h, p, m)
    sv.apply_c1q(q, p, np.diag([1, f]))
    oracle.apply_c1q(h, q, p, np.diag([1, f]))
    for d in (np.array([1, 1, 1, f]), np.array([1, f, 1, 1]), np.array([f, 1, 1, 1])):
        sv.apply_diag2(p, q, d)
        oracle.apply_diag(h, p, q, d)
    assert np.abs(sv.download() - h).max() < 4 * TOL[dt]


def test_tiled_toffoli_and_scale(dt, big_states):
    h = big_states[dt].copy()
    sv = on_gpu(h, dt)
    sv.apply_mc1q([0, 16], 8, [[0, 1], [1, 0]])
    j = np.arange(1 << NT)
    sel = ((j & 1) == 1) & (((j >> 16) & 1) == 1)
    want = h.copy()
    want[sel] = h[sel ^ (1 << 8) * 0 + 0][...] if False else h[np.where(sel, j ^ (1 << 8), j)][sel]
    assert np.array_equal(sv.download(), want)
    sv.scale(0.5 + 0.5j)
    assert np.abs(sv.download() - (0.5 + 0.5j) * want.astype(complex)).max() < TOL[dt]


@pytest.mark.parametrize("controls,t", [([7, 16], 2), ([4], 2), ([9, 11, 15], 0), ([2, 13], 6), ([3], 10)])
def test_tiled_multi_controlled_dispatch(dt, big_states, controls, t):
    h = big_states[dt].copy()
    m = haar_unitary(2, SplitMix64(sum(controls) * 7 + t))
    sv = on_gpu(h, dt)
    sv.apply_mc1q(controls, t, m)
    j = np.arange(1 << NT)
    cmask = sum(1 << c for c in controls)
    sel = ((j & cmask) == cmask) & (((j >> t) & 1) == 0)
    a, b = h[sel].astype(complex), h[j[sel] | (1 << t)].astype(complex)
    want = h.astype(complex).copy()
    want[sel] = m[0, 0] * a + m[0, 1] * b
    want[j[sel] | (1 << t)] = m[1, 0] * a + m[1, 1] * b
    assert np.abs(sv.download() - want).max() < TOL[dt]
    sv.upload(h)
    sv.apply_mc1q(controls, t, [[0, 1], [1, 0]])
    want = h.copy()
    want[sel], want[j[sel] | (1 << t)] = h[j[sel] | (1 << t)], h[sel]
    assert np.array_equal(sv.download(), want)


@pytest.mark.parametrize("positions", [[0, 1, 2, 3, 4], [4, 3, 2, 1, 0], [13, 2, 7, 0, 10], [9, 10, 11, 12, 13],
                                       [1, 5, 6, 11, 3]])
def test_block5_mfma_vs_valu_and_gather(dt, positions):
    from iqs_amd._lib import call
    n = 14
    U = haar_unitary(32, SplitMix64(sum(p * (i + 3) for i, p in enumerate(positions))))
    h = rand_state(n, dt, 41)
    sv = on_gpu(h, dt)
    sv.apply_kq(positions, U)  # matrix cores, three real products
    got = sv.download()
    sv.upload(h)
    call("iqsh_set_block_mfma", 2)
    try:
        sv.apply_kq(positions, U)  # matrix cores, one 64 x 64 real product
        got4 = sv.download()
        sv.upload(h)
        call("iqsh_set_block_mfma", 0)
        sv.apply_kq(positions, U)  # vector ALUs
    finally:
        call("iqsh_set_block_mfma", 1)
    valu = sv.download()
    a = h.astype(np.complex128)
    want = a.copy()
    j = np.arange(1 << n)
    mask = sum(1 << p for p in positions)
    bases = j[(j & mask) == 0]
    idx = np.stack([bases | sum(((r >> b) & 1) << positions[b] for b in range(5)) for r in range(32)])
    want[idx] = U @ a[idx]
    t = 1e-12 if dt == "c128" else 3e-5
    assert np.abs(got - want).max() < t
    assert np.abs(got4 - want).max() < t
    assert np.abs(valu - want).max() < t


def test_block_kernel_mode_is_checked():
    from iqs_amd._lib import IqsError, call
    for bad in (-1, 3):
        with pytest.raises(IqsError):
            call("iqsh_set_block_mfma", bad)


@pytest.mark.parametrize("positions", [[0, 1, 2], [13, 2, 7], [12, 13, 11], [1, 3, 5, 7], [12, 13, 0, 6], [3, 2, 1, 0]])
def test_block34_padded_to_mfma(dt, positions):
    """3- and 4-qubit blocks run as U (x) I on the matrix cores (padded with the highest
    free positions): compare with the vector-ALU kernel and an explicit gather."""
    from iqs_amd._lib import call
    n, k = 14, len(positions)
    U = haar_unitary(1 << k, SplitMix64(7 * sum(positions) + k))
    h = rand_state(n, dt, 43)
    sv = on_gpu(h, dt)
    sv.apply_kq(positions, U)
    got = sv.download()
    sv.upload(h)
    call("iqsh_set_block_mfma", 0)
    try:
        sv.apply_kq(positions, U)
    finally:
        call("iqsh_set_block_mfma", 1)
    valu = sv.download()
    a = h.astype(np.complex128)
    want = a.copy()
    j = np.arange(1 << n)
    mask = sum(1 << p for p in positions)
    bases = j[(j & mask) == 0]
    idx = np.stack([bases | sum(((r >> b) & 1) << positions[b] for b in range(k)) for r in range(1 << k)])
    want[idx] = U @ a[idx]
    t = 1e-12 if dt == "c128" else 3e-5
    assert np.abs(got - want).max() < t
    assert np.abs(valu - want).max() < t


# ---- registers of the tiled size (n >= 17): lane-shuffle 4x4 and line-aligned controls ----

@pytest.mark.parametrize("ph,pl", [(0, 1), (1, 0), (2, 16), (16, 0), (5, 3), (3, 5), (7, 4), (6, 5), (12, 9)])
def test_2q_dense_tiled_low_positions(dt, ph, pl):
    n = 17
    m = haar_unitary(4, SplitMix64(31 * ph + pl))
    h = rand_state(n, dt, ph + 7 * pl)
    sv = on_gpu(h, dt)
    sv.apply_2q(ph, pl, m)
    oracle.apply_2q(h, ph, pl, m)
    assert np.abs(sv.download() - h).max() < TOL[dt]


@pytest.mark.parametrize("p1,p2", [(0, 16), (16, 0), (1, 2), (2, 1), (0, 5), (4, 9), (3, 0)])
def test_swaplike_tiled_low_positions(dt, p1, p2):
    n = 17
    h = rand_state(n, dt, 11 + p1)
    sv = on_gpu(h, dt)
    th = 0.41 + p2
    m = np.array([[math.cos(th), 1j * math.sin(th)], [1j * math.sin(th), math.cos(th)]])
    sv.apply_swaplike(p1, p2, m)
    oracle.apply_swaplike(h, min(p1, p2), max(p1, p2), m)
    assert np.abs(sv.download() - h).max() < TOL[dt]
    sv.upload(h)  # identical starting bits for the exact-move check
    sv.apply_swaplike(p1, p2, [[0, 1], [1, 0]])  # exact move
    oracle.apply_swaplike(h, min(p1, p2), max(p1, p2), np.array([[0, 1], [1, 0]]))
    assert np.array_equal(sv.download(), h)


@pytest.mark.parametrize("c,t", [(3, 2), (4, 0), (5, 1), (5, 4), (3, 5), (4, 3)])
def test_controlled_tiled_line_aligned_control(dt, c, t):
    n = 17
    h = rand_state(n, dt, 3 * c + t)
    sv = on_gpu(h, dt)
    m = random_su2(SplitMix64(c * 100 + t))
    sv.apply_c1q(c, t, m)
    oracle.apply_c1q(h, c, t, m)
    assert np.abs(sv.download() - h).max() < TOL[dt]
    sv.upload(h)  # identical starting bits for the exact-move check
    sv.apply_c1q(c, t, [[0, 1], [1, 0]])  # CNOT: exact move
    oracle.apply_c1q(h, c, t, np.array([[0, 1], [1, 0]]))
    assert np.array_equal(sv.download(), h)
