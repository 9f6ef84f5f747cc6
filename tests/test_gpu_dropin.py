"""The intelqs_py drop-in on the GPU against the reference's own results: every golden
fixture (written by the compiled reference) replayed through QubitRegister /
QubitRegisterSP; fused execution against unfused; closed forms from the reference's unit
tests (unit_test/include/*.hpp) re-expressed through the Python module."""
import glob
import math
import os

import numpy as np
import pytest

from iqs_amd import runner
from iqs_amd.circuits import Circuit, brick_circuit, random_1q2q_circuit

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = sorted(os.path.basename(p)[:-5] for p in glob.glob(os.path.join(GOLDEN, "*.circ")))


def tol(prec):
    return 1e-12 if prec == "dp" else 1e-5


def load(name):
    text = open(os.path.join(GOLDEN, name + ".circ")).read()
    res = [l for l in open(os.path.join(GOLDEN, name + ".res")).read().splitlines() if l]
    return Circuit.parse(text), res, np.load(os.path.join(GOLDEN, name + ".npz"))


def compare_results(got, want, prec):
    g, w = runner.parse_results(got), runner.parse_results(want)
    assert [x[0] for x in g] == [x[0] for x in w]
    # SP reductions: the reference accumulates in float serially, the GPU in FP64.
    t = 1e-11 if prec == "dp" else 5e-5
    for (name, gv), (_, wv) in zip(g, w):
        assert len(gv) == len(wv), name
        assert np.abs(np.array(gv) - np.array(wv)).max() < t, (name, gv, wv)


@pytest.mark.parametrize("name", CASES)
def test_fixture_through_dropin(name):
    c, want_res, g = load(name)
    reg, res = runner.run(c)
    st = runner.state_of(reg)
    if "state" in g.files:
        assert np.abs(st - g["state"]).max() < tol(c.precision)
    else:
        assert np.abs(st[g["idx"]] - g["amp"]).max() < tol(c.precision)
        a = st.astype(np.complex128)
        assert abs(np.vdot(a, a).real - g["norm2"][0]) < 1e-11
    compare_results(res, want_res, c.precision)


@pytest.mark.parametrize("name", ["brick10_dp", "brick10_sp", "random1q2q_dp", "random1q2q_sp", "gates2q_dp"])
@pytest.mark.parametrize("k", [1, 2, 3, 5])
def test_fusion_matches_unfused_reference(name, k):
    c, _, g = load(name)
    reg = runner.new_register(c)
    reg.SetFusionBlockQubits(k)
    reg.TurnOnFusion()
    reg, _ = runner.run(c, reg)
    st = runner.state_of(reg)
    assert np.abs(st - g["state"]).max() < (1e-11 if c.precision == "dp" else 2e-5)


def test_fused_brick_on_tiled_sizes():
    """Fusion on a register large enough for the tiled kernels (n = 16)."""
    c = brick_circuit(16, 4, 3)
    ref, _ = runner.run(c)
    reg = runner.new_register(c)
    reg.TurnOnFusion()
    reg, _ = runner.run(c, reg)
    assert np.abs(runner.state_of(reg) - runner.state_of(ref)).max() < 1e-12
    c = random_1q2q_circuit(17, 40, 9, precision="sp")
    ref, _ = runner.run(c)
    reg = runner.new_register(c)
    reg.TurnOnFusion()
    reg, _ = runner.run(c, reg)
    assert np.abs(runner.state_of(reg) - runner.state_of(ref)).max() < 2e-5


# ---- closed forms from the reference's unit tests ----
import intelqs_py as iqs  # noqa: E402


def test_hadamard_and_rotations_closed_form():
    """apply_1q_gate_test.hpp:48-180 (4 qubits, gate on qubit 3: |0> -> index 0, |1> -> 8)."""
    s2 = math.sqrt(2.0)
    psi = iqs.QubitRegister(4, "base", 0, 0)
    psi.ApplyHadamard(3)
    assert abs(psi.ComputeNorm() - 1.0) < 1e-15
    assert abs(psi.GetGlobalAmplitude(0) - 1 / s2) < 1e-15 and abs(psi.GetGlobalAmplitude(8) - 1 / s2) < 1e-15
    psi.Initialize("base", 8)
    psi.ApplyHadamard(3)
    assert abs(psi.GetGlobalAmplitude(8) + 1 / s2) < 1e-15
    psi.Initialize("base", 0)
    psi.ApplyRotationX(3, 0.83)
    assert abs(psi.GetGlobalAmplitude(0) - math.cos(0.83 / 2)) < 1e-15
    assert abs(psi.GetGlobalAmplitude(8) - complex(0, -math.sin(0.83 / 2))) < 1e-15


def test_rxy_decomposition_on_random_state():
    """apply_1q_gate_test.hpp:188-234: RXY(phi, theta) = RZ(phi) RX(theta) RZ(-phi)."""
    rng = iqs.RandomNumberGenerator()
    rng.SetSeedStreamPtrs(777)
    a = iqs.QubitRegister(4, "base", 0, 0)
    a.SetRngPtr(rng)
    a.Initialize("rand", 0)
    assert abs(a.ComputeNorm() - 1) < 1e-14
    b = iqs.QubitRegister(a)
    phi, theta = 0.62, 1.37
    a.ApplyRotationXY(2, phi, theta)
    b.ApplyRotationZ(2, -phi)
    b.ApplyRotationX(2, theta)
    b.ApplyRotationZ(2, phi)
    assert a.MaxAbsDiff(b) < 1e-14


def test_swap_exact_and_three_cnots():
    """apply_swap_gate_test.hpp:65-247."""
    psi = iqs.QubitRegister(3, "base", 0, 0)
    for i in range(8):
        psi[i] = complex(i, 0)
    psi.ApplySwap(0, 1)
    assert [psi[i].real for i in range(8)] == [0, 2, 1, 3, 4, 6, 5, 7]
    a = iqs.QubitRegister(12, "++++", 0, 0)
    for q in range(12):
        a.ApplyRotationY(q, 0.1 * q + 0.05)
    b = iqs.QubitRegister(a)
    a.ApplySwap(2, 9)
    b.ApplyCPauliX(2, 9)
    b.ApplyCPauliX(9, 2)
    b.ApplyCPauliX(2, 9)
    assert a.MaxAbsDiff(b) == 0.0


def test_expectations_and_probabilities():
    """expectation_values_test.hpp:40-238, measure_test.hpp:37-57."""
    psi = iqs.QubitRegister(3, "base", 0, 0)
    psi.ApplyPauliX(1)
    psi.ApplyHadamard(2)
    assert psi.GetProbability(0) == 0.0
    assert abs(psi.GetProbability(1) - 1.0) < 1e-15
    assert abs(psi.GetProbability(2) - 0.5) < 1e-15
    assert abs(psi.ExpectationValueZ(1) + 1.0) < 1e-14
    assert abs(psi.ExpectationValueX(2) - 1.0) < 1e-14
    assert abs(psi.ExpectationValueY(0)) < 1e-14
    assert abs(psi.ExpectationValue([1, 2], [3, 1], 2.0) + 2.0) < 1e-14
    before = np.array(psi, copy=True)
    psi.ExpectationValue([0, 1, 2], [2, 2, 1], 1.0)
    assert np.array_equal(np.array(psi, copy=True), before)  # read-only pass


def test_permutation_placement_and_amplitudes():
    """qureg_permute_test.hpp / permutation_test.hpp: values follow the program index."""
    n = 6
    psi = iqs.QubitRegister(n, "base", 0, 0)
    for i in range(1 << n):
        psi.SetGlobalAmplitude(i, complex(i, -i))
    psi.PermuteQubits([1, 0, 2, 5, 4, 3], "direct")
    assert psi.GetQubitPermutation() == [1, 0, 2, 5, 4, 3]
    for i in range(1 << n):
        assert psi.GetGlobalAmplitude(i) == complex(i, -i)
    raw = np.array(psi, copy=True)
    # data index: bit map[q] carries program bit q
    for i in (1, 2, 8, 33):
        d = sum(((i >> q) & 1) << [1, 0, 2, 5, 4, 3][q] for q in range(n))
        assert raw[d] == complex(i, -i)
    psi.EmulateSwap(0, 5)  # relabels positions: map[0] <-> map[5], no data moves
    newmap = [3, 0, 2, 5, 4, 1]
    assert psi.GetQubitPermutation() == newmap

    def p2d(i):
        return sum(((i >> q) & 1) << newmap[q] for q in range(n))

    for i in (0, 1, 7, 40):
        assert psi.GetGlobalAmplitude(i) == raw[p2d(i)]
    psi.ApplyPauliX(0)  # flips program qubit 0, stored at position 3
    for i in (0, 1, 7, 40):
        assert psi.GetGlobalAmplitude(i) == raw[p2d(i ^ 1)]


def test_qaoa_ring_and_weighted_cuts():
    """qaoa_features_test.hpp:39-150."""
    adj = [0, 1, 0, 0, 0, 1, 1, 0, 1, 0, 0, 0, 0, 1, 0, 1, 0, 0, 0, 0, 1, 0, 1, 0, 0, 0, 0, 1, 0, 1, 1, 0, 0, 0, 1, 0]
    diag = iqs.QubitRegister(6, "base", 0, 0)
    mx = iqs.InitializeVectorAsMaxCutCostFunction(diag, adj)
    assert mx == 6
    assert diag.GetGlobalAmplitude(0) == 0 and diag.GetGlobalAmplitude(63) == 0
    vals = np.array(diag, copy=True).real
    assert np.all(np.abs(vals - 1.0) > 1e-15)
    psi = iqs.QubitRegister(6, "++++", 0, 0)
    iqs.ImplementQaoaLayerBasedOnCostFunction(psi, diag, 0.4)
    for q in range(6):
        psi.ApplyRotationX(q, 0.3)
    e = iqs.GetExpectationValueFromCostFunction(psi, diag)
    p = np.abs(np.array(psi, copy=True)) ** 2
    assert abs(e - (vals * p).sum()) < 1e-13
    w = [0, 1, 0, 0, 0, 1.4, 1, 0, 1, 0, 0, 0, 0, 1, 0, 1.4, 0, 0, 0, 0, 1.4, 0, 1, 0, 0, 0, 0, 1, 0, 1, 1.4, 0, 0, 0, 1, 0]
    dw = iqs.QubitRegister(6, "base", 0, 0)
    iqs.InitializeVectorAsWeightedMaxCutCostFunction(dw, w)
    assert abs(dw.GetGlobalAmplitude(4) - 2.4) < 1e-15
    assert abs(dw.GetGlobalAmplitude(36) - 4.8) < 1e-15


@pytest.mark.parametrize("cls", ["QubitRegister", "QubitRegisterSP"])
def test_cost_register_of_dense_graphs(cls):
    """More edges than one 64-bit edge mask: the complete graph K_13 (78 edges) has cut
    k (13 - k) for a basis index with k ones (max 42), and a weighted dense graph (random
    weights, 70 edges) against the brute-force cut of every index."""
    n = 13
    reg = getattr(iqs, cls)
    diag = reg(n, "base", 0, 0)
    adj = [0 if u == v else 1 for v in range(n) for u in range(n)]
    assert iqs.InitializeVectorAsMaxCutCostFunction(diag, adj) == 42
    vals = np.array(diag, copy=True).real
    k = np.array([bin(i).count("1") for i in range(1 << n)])
    assert np.array_equal(vals, (k * (n - k)).astype(vals.dtype))
    r = np.random.default_rng(13)
    w = np.zeros((n, n))
    edges = [(a, b) for a in range(n) for b in range(a + 1, n)]
    for e in r.permutation(len(edges))[:70]:
        a, b = edges[e]
        w[a, b] = w[b, a] = round(float(r.uniform(0.5, 2.0)), 3)
    dw = reg(n, "base", 0, 0)
    mx = iqs.InitializeVectorAsWeightedMaxCutCostFunction(dw, [float(x) for x in w.ravel()])
    idx = np.arange(1 << n)
    bits = (idx[:, None] >> np.arange(n)) & 1
    want = np.einsum("ia,ab,ib->i", bits, w, 1 - bits)  # sum of w over edges with different bits
    got = np.array(dw, copy=True).real
    tol = 1e-12 if cls == "QubitRegister" else 1e-4
    assert np.abs(got - want).max() < tol
    assert abs(mx - want.max()) < tol


def test_gate_counter_and_toffoli_truth_table():
    """gate_counter_test.hpp; Toffoli on basis states."""
    psi = iqs.QubitRegister(4, "base", 0, 0)
    psi.EnableStatistics()
    psi.ApplyHadamard(0)
    psi.ApplyCPauliX(0, 1)
    psi.ApplyToffoli(0, 1, 2)
    one, two, total, depth = psi.GetGateCounts()
    assert (one, two, total) == (1, 6, 7)
    for j in range(8):
        r = iqs.QubitRegister(3, "base", j, 0)
        r.ApplyToffoli(0, 1, 2)
        want = j ^ 4 if (j & 3) == 3 else j
        assert r.GetGlobalAmplitude(want) == 1


def test_classical_bits_collapse_overlap_buffer():
    psi = iqs.QubitRegister(5, "base", 0b10110, 0)
    assert psi.IsClassicalBit(1) and psi.GetClassicalValue(1) and not psi.GetClassicalValue(0)
    psi.ApplyHadamard(0)
    assert not psi.IsClassicalBit(0)
    psi.CollapseQubit(0, True)
    psi.Normalize()
    assert psi.GetClassicalValue(0)
    phi = iqs.QubitRegister(psi)
    assert abs(psi.ComputeOverlap(phi) - 1) < 1e-15
    view = np.array(phi, copy=False)  # host mirror, written back on the next GPU operation
    view[:] = 0
    view[3] = 1j
    assert phi.ComputeNorm() == 1.0
    assert phi.GetGlobalAmplitude(3) == 1j


def test_sp_register_and_errors():
    psi = iqs.QubitRegisterSP(10, "++++", 0, 0)
    for q in range(10):
        psi.ApplyRotationZ(q, 0.2 * q)
    assert abs(psi.ComputeNorm() - 1) < 1e-6
    with pytest.raises((IndexError, RuntimeError, ValueError)):
        psi.ApplyHadamard(10)
    with pytest.raises(RuntimeError, match="2x2"):
        psi.Apply1QubitGate(0, np.eye(3, dtype=np.complex64))


def test_qubit_register_metric_counts_and_depth():
    """QubitRegisterMetric (include/QubitRegisterMetric.hpp of the reference): 1q gates add a
    layer on their qubit, 2q gates lift both qubits one layer above the deeper one; the
    state is the plain register's."""
    m = iqs.QubitRegisterMetric(5)
    m.Initialize("base", 0)
    p = iqs.QubitRegister(5, "base", 0, 0)
    ops = [("ApplyHadamard", 0), ("ApplyHadamard", 1), ("ApplyRotationX", 1, 0.3), ("ApplyCPauliX", 0, 1),
           ("ApplyRotationZ", 4, 1.1), ("ApplyCPauliX", 1, 2), ("ApplyRotationY", 3, 0.5)]
    for op in ops:
        getattr(m, op[0])(*op[1:])
        getattr(p, op[0])(*op[1:])
    G = np.array([[0, 1j], [1j, 0]])
    m.ApplyControlled1QubitGate(2, 3, G)
    p.ApplyControlled1QubitGate(2, 3, G)
    # depths: q0 1, q1 2 -> CX(0,1): 3,3 -> CX(1,2): 4,4 -> q3 1 -> C(2,3): 5,5
    assert (m.GetOneQubitGateCount(), m.GetTwoQubitGateCount(), m.GetTotalQubitGateCount()) == (5, 3, 8)
    assert m.GetParallelDepth() == 5
    assert np.array_equal(np.array(m, copy=True), np.array(p, copy=True))


# ---- sub-range gate helpers (src/qureg_apply1qubitgate.cpp:173-230,
#      src/qureg_applyctrl1qubitgate.cpp:228-346) ----

def _loop_nest(s, c11, c12, c13, c21, c22, c23, c31, c32, shift, m):
    """Loop_TN's nest (src/highperfkernels.cpp:397-486) as the reference runs it."""
    s = s.copy()
    for i in range(c11, c12, c13):
        for j in range(i + c21, i + c22, c23):
            for k in range(j + c31, j + c32):
                a, b = s[k], s[k + shift]
                s[k], s[k + shift] = m[0, 0] * a + m[0, 1] * b, m[1, 0] * a + m[1, 1] * b
    return s


def _ref_1q_helper(s, P, sind, eind, m):
    return _loop_nest(s, sind, eind, 2 << P, 0, 1, 1, 0, 1 << P, 1 << P, m)


def _ref_c1q_helper(s, C, T, sind, eind, m, log2llc):
    if C > T and C >= log2llc and s.size > eind - sind:
        return _ref_1q_helper(s, T, sind, eind, m) if (sind >> C) & 1 else s
    if C > T:
        return _loop_nest(s, sind, eind, 2 << C, 1 << C, 2 << C, 2 << T, 0, 1 << T, 1 << T, m)
    return _loop_nest(s, sind, eind, 2 << T, 0, 1 << T, 2 << C, 1 << C, 2 << C, 1 << T, m)


@pytest.mark.parametrize("cls", ["QubitRegister", "QubitRegisterSP"])
def test_sub_range_helpers(cls):
    n = 10
    dt = np.complex128 if cls == "QubitRegister" else np.complex64
    tol = 1e-13 if cls == "QubitRegister" else 2e-6
    r = np.random.default_rng(3)
    h = (r.normal(size=1 << n) + 1j * r.normal(size=1 << n)).astype(dt)
    h /= np.linalg.norm(h)
    m = np.array([[0.6, 0.8j], [0.8j, 0.6]], dtype=np.complex128)

    def fresh():
        reg = getattr(iqs, cls)(n, "base", 0, 0)
        np.array(reg, copy=False)[:] = h
        return reg

    cases_1q = [(0, 0, 64), (3, 128, 512), (6, 256, 1024), (2, 40, 200), (9, 0, 1024), (4, 992, 1024)]
    for P, sind, eind in cases_1q:
        reg = fresh()
        assert reg.Apply1QubitGate_helper(P, m.astype(dt), sind, eind)
        want = _ref_1q_helper(h.astype(np.complex128), P, sind, eind, m)
        assert np.abs(np.array(reg, copy=True) - want).max() < tol, (P, sind, eind)
    # control above the cache block (log2llc = 5 after TurnOnFusion(5)), below it, and C < T
    for C, T, sind, eind in [(8, 2, 256, 288), (8, 2, 0, 32), (7, 3, 128, 256), (3, 6, 0, 512), (1, 0, 100, 300),
                             (9, 4, 0, 1024)]:
        reg = fresh()
        reg.TurnOnFusion(5)
        reg.TurnOffFusion()  # keeps log2llc = 5, fusion off
        reg.ApplyControlled1QubitGate_helper(C, T, m.astype(dt), sind, eind)
        want = _ref_c1q_helper(h.astype(np.complex128), C, T, sind, eind, m, 5)
        assert np.abs(np.array(reg, copy=True) - want).max() < tol, (C, T, sind, eind)
    # SpecializeV2 constants on a sub-range: RX(0.7) through the spec path equals the matrix
    reg = fresh()
    reg.TurnOnSpecializeV2()
    reg.Apply1QubitGate_helper(5, np.eye(2, dtype=dt), 64, 512, 1, 0.7)  # GateSpec1Q::RotationX
    c, s_ = np.cos(0.35), np.sin(0.35)
    rx = np.array([[c, -1j * s_], [-1j * s_, c]])
    assert np.abs(np.array(reg, copy=True) - _ref_1q_helper(h.astype(np.complex128), 5, 64, 512, rx)).max() < tol
    with pytest.raises(Exception):
        fresh().Apply1QubitGate_helper(3, m.astype(dt), 0, 2048)  # beyond the register
