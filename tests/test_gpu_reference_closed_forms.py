"""The reference's own unit-test closed forms, through the intelqs_py drop-in on the GPU, at
the reference's sizes and at 26-30 qubits (two-dimensional grids, positions up to 29).

- state_initialization_test.hpp:207-248  basis states under a qubit permutation
- utility_methods_test.hpp:36-160          ComputeOverlap / MaxAbsDiff / MaxL2NormDiff,
                                           AmplitudeWiseScalarMultiplication / Sum
- single_qubit_gates_test.hpp:43-120       Pauli, rotation and custom-gate probabilities
- chunking_communication_test.hpp:45-161   H on every qubit = |+...+>, custom gate and
                                           CNOT ladders on the highest six qubits
- examples/quantum_fourier_transform.cpp:28-73,130-170  QFT circuit against the DFT

Where a reference test compares registers that on one GPU are computed identically (the
MPI chunk sizes of chunking_communication_test.hpp), the same circuit is also checked
against an independent answer: the oracle on the whole state, or the cluster-product
closed form of tests/_factored.py.  Tolerances are the reference's (1e-15 / 1e-12 on
values of order one) except where stated with the reason.
"""
import cmath
import math

import numpy as np
import pytest

import oracle
from _factored import Factored
from iqs_amd.circuits import SplitMix64, haar_unitary

pytestmark = pytest.mark.gpu

S2 = 1.0 / math.sqrt(2.0)
# G of single_qubit_gates_test.hpp:29-32 and chunking_communication_test.hpp:122-125
GMAT = np.array([[0.592056606032915 + 0.459533060553574j, -0.314948020757856 - 0.582328159830658j],
                 [0.658235557641767 + 0.070882241549507j, 0.649564427121402 + 0.373855203932477j]])


def m():
    import intelqs_py
    return intelqs_py


def deq(a, b, ulps=4):
    """gtest's ASSERT_DOUBLE_EQ: within 4 units in the last place."""
    return abs(a - b) <= ulps * np.spacing(max(abs(a), abs(b), np.finfo(float).tiny))


def p2d(i, perm):
    """program index -> data index under the map (bit map[q] of the data index carries
    program bit q: Permutation::program2data_, include/permutation.hpp:249-264)."""
    return sum(((i >> q) & 1) << perm[q] for q in range(len(perm)))


@pytest.mark.parametrize("n", [10, 28])
def test_basis_state_under_qubit_permutation(n):
    iqs = m()
    index = 1 + 4 + 32 + 256 + 512
    psi = iqs.QubitRegister(n, "base", index, 0)
    assert psi.GetGlobalAmplitude(index) == 1
    assert psi[index] == 1
    perm = [0, 1, 2, 3, 4, 5, 9, 7, 8, 6] + list(range(10, n))
    if n > 10:  # also move program qubits 2 and n-1 between the bottom and the top position
        perm[2], perm[n - 1] = n - 1, 2
    psi.PermuteQubits(perm, "direct")
    assert psi.GetQubitPermutation() == perm
    assert psi.GetGlobalAmplitude(index) == 1
    assert psi[p2d(index, perm)] == 1
    assert psi[index] == (1 if p2d(index, perm) == index else 0)
    index = 1 + 64 + ((1 << (n - 1)) if n > 10 else 0)
    psi.Initialize("base", index)
    assert psi.GetGlobalAmplitude(index) == 1
    assert psi[p2d(index, perm)] == 1
    assert psi.ComputeNorm() == 1.0
    assert psi.GetProbability(6) == 1.0 and psi.GetProbability(9) == 0.0


@pytest.mark.parametrize("n", [4, 26])
def test_difference_of_two_states(n):
    """utility_methods_test.hpp:36-103."""
    iqs = m()
    a = iqs.QubitRegister(n, "base", 0, 0)
    b = iqs.QubitRegister(n, "base", 4, 0)
    assert deq(abs(a.ComputeOverlap(a)) ** 2, 1.0)
    assert deq(abs(b.ComputeOverlap(a)) ** 2, 0.0)
    assert deq(b.MaxAbsDiff(a), 1.0)
    assert deq(b.MaxL2NormDiff(a), 2.0)
    b.ApplyHadamard(2)
    assert deq(abs(b.ComputeOverlap(a)) ** 2, 0.5)
    assert deq(b.ComputeOverlap(a).real, S2)
    assert deq(b.MaxAbsDiff(a), S2)
    assert deq(b.MaxL2NormDiff(a), (1 - S2) * (1 - S2) + 0.5)
    a.ApplyPauliX(2)
    assert deq(b.ComputeOverlap(a).real, -S2)
    a.Initialize("base", 12)
    b.Initialize("base", 12)
    b.ApplyPauliZ(3)
    assert deq(b.ComputeOverlap(a).real, -1.0)
    assert deq(b.MaxL2NormDiff(a), 4.0)
    for r in (a, b):
        r.ApplyHadamard(2)
        r.ApplyHadamard(3)
    assert deq(b.MaxL2NormDiff(a), 4.0)


@pytest.mark.parametrize("n", [4, 22])
def test_amplitude_wise_multiplication_and_sum(n):
    """utility_methods_test.hpp:120-157 (random states, seed 7777); the larger register
    is checked on sampled amplitudes."""
    iqs = m()
    rng = iqs.RandomNumberGenerator()
    rng.SetSeedStreamPtrs(7777)
    psi = iqs.QubitRegister(n, "base", 0, 0)
    psi.SetRngPtr(rng)
    psi.Initialize("rand", 1)
    phi = iqs.QubitRegister(n, "base", 0, 0)
    phi.SetRngPtr(rng)
    phi.Initialize("rand", 1)
    assert psi[0] != phi[0] and psi[11] != phi[1]
    idx = range(1 << n) if n <= 10 else np.random.default_rng(n).integers(0, 1 << n, 512)
    copy = iqs.QubitRegister(psi)
    f = complex(0.4, 1.1)
    psi.AmplitudeWiseScalarMultiplication(f)
    for i in idx:
        assert abs(psi[int(i)] - copy[int(i)] * f) < 1e-15
    psi2 = iqs.QubitRegister(copy)
    f = complex(-0.2, 0.8)
    psi2.AmplitudeWiseSum(phi, f)
    for i in idx:
        assert abs(psi2[int(i)] - (copy[int(i)] + phi[int(i)] * f)) < 1e-15


@pytest.mark.parametrize("n", [10, 30])
def test_pauli_operators(n):
    """single_qubit_gates_test.hpp:43-79, with the reference's qubit num_qubits-1 = 9 at
    the top position n-1."""
    iqs = m()
    top = 1 << (n - 1)
    psi = iqs.QubitRegister(n, "base", 1 + 4 + 32 + top, 0)
    assert deq(psi.ComputeNorm(), 1.0)
    assert deq(psi.GetProbability(0), 1.0) and deq(psi.GetProbability(2), 1.0)
    assert deq(psi.GetProbability(n - 2), 0.0) and deq(psi.GetProbability(n - 1), 1.0)
    psi.ApplyPauliX(5)
    psi.ApplyPauliX(6)
    assert deq(psi.GetProbability(5), 0.0) and deq(psi.GetProbability(6), 1.0)
    idx = 2 + 4 + (1 << (n - 4))
    psi.Initialize("base", idx)
    assert deq(psi.ComputeNorm(), 1.0)
    assert deq(psi.GetProbability(2), 1.0) and deq(psi.GetProbability(n - 4), 1.0)
    psi.ApplyPauliZ(1)
    assert deq(psi.GetGlobalAmplitude(idx).real, -1.0)
    psi.ApplyPauliZ(2)
    assert deq(psi.GetGlobalAmplitude(idx).real, 1.0)
    psi.ApplyPauliZ(3)
    assert deq(psi.GetGlobalAmplitude(idx).real, 1.0)
    psi.ApplyPauliX(4)
    assert deq(psi.GetGlobalAmplitude(idx + 16).real, 1.0) and psi.GetGlobalAmplitude(idx + 16).imag == 0.0


@pytest.mark.parametrize("n", [10, 30])
def test_rotations(n):
    """single_qubit_gates_test.hpp:83-106."""
    iqs = m()
    idx = 1 + 4 + 32 + (1 << (n - 1))
    psi = iqs.QubitRegister(n, "base", idx, 0)
    psi.ApplyRotationX(1, math.pi)
    psi.ApplyPauliX(1)
    assert abs(psi.GetProbability(1)) < 1e-15
    assert deq(psi.GetGlobalAmplitude(idx).real, 0.0) and deq(psi.GetGlobalAmplitude(idx).imag, -1.0)
    # The reference asserts these to 4 ULP (ASSERT_DOUBLE_EQ); with sums of several
    # rounded products (and 1/sqrt(2) squared twice for the last one) 1e-15 is used.
    ang = math.pi / 2.34
    psi.ApplyRotationX(0, ang)
    assert abs(psi.GetProbability(0) - math.cos(ang / 2) ** 2) < 1e-15
    psi.ApplyRotationX(n - 2, ang)
    assert abs(psi.GetProbability(n - 2) - math.sin(ang / 2) ** 2) < 1e-15
    ang = math.pi * 0.29
    psi.ApplyHadamard(5)
    psi.ApplyRotationZ(5, ang)
    psi.ApplyHadamard(5)
    assert abs(psi.GetProbability(5) - math.cos(ang / 2) ** 2) < 1e-15


@pytest.mark.parametrize("n", [10, 30])
def test_custom_gate(n):
    """single_qubit_gates_test.hpp:110-120: G on every qubit of |1+4+32+top>."""
    iqs = m()
    psi = iqs.QubitRegister(n, "base", 1 + 4 + 32 + (1 << (n - 1)), 0)
    for q in range(n):
        psi.Apply1QubitGate(q, GMAT)
    assert deq(psi.GetProbability(0), psi.GetProbability(2))
    assert deq(psi.GetProbability(4), psi.GetProbability(6))
    assert abs(psi.GetProbability(5) - (1 - psi.GetProbability(8))) < 1e-14
    # and each probability is the single-qubit closed form
    g = np.asarray(GMAT)
    for q, bit in ((0, 1), (4, 0), (n - 1, 1), (n - 2, 0)):
        assert abs(psi.GetProbability(q) - abs(g[1, bit]) ** 2) < 1e-14


@pytest.mark.parametrize("n", [14, 28])
def test_chunking_initialization_and_hadamards(n):
    """chunking_communication_test.hpp:45-100 (the temporary-space sizes of the MPI build
    are accepted and do not change the result on one GPU)."""
    iqs = m()
    local = 1 << n
    regs = [iqs.QubitRegister(n, "base", i, t) for i, t in ((1, local // 2), (10200, local // 4), (5003, local // 8))]
    for r, i in zip(regs, (1, 10200, 5003)):
        assert abs(r.ComputeNorm() - 1) <= 1e-15
        assert deq(r.GetGlobalAmplitude(i).real, 1.0)
    h = regs[0]
    h.ApplyPauliX(0)
    for q in range(n):
        h.ApplyHadamard(q)
    plus = iqs.QubitRegister(n, "++++", 0, 0)
    # Every amplitude after the Hadamards is the same product of n factors
    # 1/sqrt(2) = 0.70710678118654746 (7.5e-17 below the exact value) with n roundings, so
    # sum |a|^2 = 1 - 2n * ~1e-16; the FP64 tree sum here adds the equal terms exactly (the
    # reference's serial sum rounds its way to 1 - 1.1e-16 at n = 14 and passes 1e-15).
    tol = 2 * n * 1.2e-16
    assert abs(h.ComputeOverlap(plus) - 1) <= tol
    regs = [iqs.QubitRegister(n, "base", 0, t) for t in (local // 2, local // 4, local // 8)]
    for r in regs:
        for q in range(n):
            r.ApplyHadamard(q)
    for a, b in ((0, 1), (0, 2), (2, 1)):
        assert abs(regs[a].ComputeOverlap(regs[b]) - 1) <= tol
    assert abs(regs[0].ComputeOverlap(plus) - 1) <= tol


@pytest.mark.parametrize("n", [14, 28])
def test_chunking_custom_gate_and_cnot_ladders(n):
    """chunking_communication_test.hpp:104-161: G on every qubit, and CNOT on every ordered
    pair of the six highest qubits (both directions).  From |0> the CNOT sequence is the
    identity, so the same sequence is also applied to a product of random one-qubit
    states, against the oracle on the whole state (n = 14) or the cluster-product closed
    form (n = 28)."""
    iqs = m()
    local = 1 << n
    regs = [iqs.QubitRegister(n, "base", 0, t) for t in (local // 2, local // 4, local // 8)]
    for r in regs:
        for q in range(n):
            r.Apply1QubitGate(q, GMAT)
    for a, b in ((0, 1), (0, 2), (2, 1)):
        assert abs(regs[a].ComputeOverlap(regs[b]) - 1) <= 1e-12
    pairs = []
    for q0 in range(n - 6, n):
        for q1 in range(q0 + 1, n):
            pairs += [(q0, q1), (q1, q0)]
    zero = iqs.QubitRegister(n, "base", 0, 0)
    for r in regs:
        for c, t in pairs:
            r.ApplyCPauliX(c, t)
    for a, b in ((0, 1), (0, 2), (2, 1)):
        assert abs(regs[a].ComputeOverlap(regs[b]) - 1) <= 1e-12
    del regs
    psi = iqs.QubitRegister(n, "base", 0, 0)
    for c, t in pairs:
        psi.ApplyCPauliX(c, t)
    assert psi.MaxAbsDiff(zero) == 0.0
    rng = SplitMix64(n)
    us = [haar_unitary(2, rng) for _ in range(n)]
    for q in range(n):
        psi.Apply1QubitGate(q, us[q])
    for c, t in pairs:
        psi.ApplyCPauliX(c, t)
    X = np.array([[0, 1], [1, 0]], dtype=complex)
    if n <= 16:
        st = np.zeros(1 << n, dtype=np.complex128)
        st[0] = 1
        for q in range(n):
            oracle.apply_1q(st, q, us[q])
        for c, t in pairs:
            oracle.apply_c1q(st, c, t, X)
        assert np.abs(np.array(psi, copy=True) - st).max() < 1e-13
    else:
        f = Factored(n)
        for q in range(n):
            f.apply_1q(q, us[q])
        for c, t in pairs:
            f.apply_c1q(c, t, X)
        idx = np.random.default_rng(n).integers(0, 1 << n, 4096, dtype=np.uint64)
        got = np.array([psi[int(i)] for i in idx])
        assert np.abs(got - f.amplitudes(idx)).max() * math.sqrt(float(1 << n)) < 1e-12


def qft(psi, n):
    """The circuit of examples/quantum_fourier_transform.cpp:28-73."""
    for i in range(n - 1, -1, -1):
        for j in range(n - 1, i, -1):
            k = j - i
            ph = np.array([[1, 0], [0, complex(math.cos(math.pi / float(1 << k)), math.sin(math.pi / float(1 << k)))]])
            psi.ApplyControlled1QubitGate(j, i, ph)
        psi.ApplyHadamard(i)
    for i in range(n // 2):
        psi.ApplySwap(i, n - 1 - i)


def test_qft_matches_dft_12_qubits():
    """examples/quantum_fourier_transform.cpp:130-170: QFT of a random state equals
    y_k = sum_j x_j e^{2 pi i jk/N} / sqrt(N) (the example's brute-force DFT)."""
    iqs = m()
    n = 12
    N = 1 << n
    r = np.random.default_rng(1)
    x = r.normal(size=N) + 1j * r.normal(size=N)
    x /= np.linalg.norm(x)
    psi = iqs.QubitRegister(n, "base", 0, 0)
    view = np.array(psi, copy=False)
    view[:] = x
    qft(psi, n)
    y = np.fft.ifft(x) * math.sqrt(N)
    assert np.abs(np.array(psi, copy=True) - y).max() < 1e-12


@pytest.mark.parametrize("x0", [0, 1, 12345678, (1 << 26) - 3])
def test_qft_of_basis_states_26_qubits(x0):
    """QFT|x0> = sum_k e^{2 pi i x0 k / N} |k> / sqrt(N) on 26 qubits, checked on sampled
    amplitudes (relative to 2^-13) and on the norm."""
    iqs = m()
    n = 26
    N = 1 << n
    psi = iqs.QubitRegister(n, "base", x0, 0)
    qft(psi, n)
    idx = np.concatenate([np.arange(64), np.random.default_rng(x0 + 1).integers(0, N, 2048)])
    got = np.array([psi[int(k)] for k in idx])
    want = np.array([cmath.exp(2j * math.pi * ((x0 * int(k)) % N) / N) for k in idx]) / math.sqrt(N)
    assert np.abs(got - want).max() * math.sqrt(N) < 1e-11
    assert abs(psi.ComputeNorm() - 1) < 1e-13
