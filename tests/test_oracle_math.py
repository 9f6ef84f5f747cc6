"""Pin the oracle's kernels against what the mathematics fixes: explicit 2^n x 2^n
operators built with Kronecker products (qubit 0 = least significant bit), closed forms
from the reference's own unit tests, a QFT against the DFT, and invariants.  A transposed
operand, a wrong bit index or a dropped term fails at least one of these."""
import math

import numpy as np
import pytest

import oracle
from iqs_amd.circuits import Circuit, SplitMix64, haar_unitary

RNG = np.random.default_rng(1234)


def rand_state(n, dt=np.complex128):
    v = RNG.normal(size=1 << n) + 1j * RNG.normal(size=1 << n)
    return (v / np.linalg.norm(v)).astype(dt)


def op_on(n, pos, m):
    """Full operator of a 1q matrix m on position pos (np.kron's first factor is the MSB)."""
    full = np.array([[1.0 + 0j]])
    for q in reversed(range(n)):
        full = np.kron(full, m if q == pos else np.eye(2))
    return full


def proj(n, pos, bit):
    return op_on(n, pos, np.diag([1.0 - bit, float(bit)]).astype(complex))


@pytest.mark.parametrize("pos", [0, 1, 3, 4])
def test_1q_matches_kronecker(pos):
    n = 5
    m = haar_unitary(2, SplitMix64(pos + 1))
    st = rand_state(n)
    want = op_on(n, pos, m) @ st
    oracle.apply_1q(st, pos, m)
    assert np.abs(st - want).max() < 1e-14


@pytest.mark.parametrize("c,t", [(0, 1), (1, 0), (4, 2), (2, 4)])
def test_c1q_matches_kronecker(c, t):
    n = 5
    m = haar_unitary(2, SplitMix64(7 + c * 5 + t))
    st = rand_state(n)
    full = proj(n, c, 0) + proj(n, c, 1) @ op_on(n, t, m)
    want = full @ st
    oracle.apply_c1q(st, c, t, m)
    assert np.abs(st - want).max() < 1e-14


def two_qubit_full(n, pa, pb, m4):
    """Operator of a 4x4 matrix with row index 2*b(pa) + b(pb)."""
    N = 1 << n
    full = np.zeros((N, N), dtype=complex)
    for col in range(N):
        cb = 2 * ((col >> pa) & 1) + ((col >> pb) & 1)
        base = col & ~((1 << pa) | (1 << pb))
        for rb in range(4):
            row = base | (((rb >> 1) & 1) << pa) | ((rb & 1) << pb)
            full[row, col] += m4[rb, cb]
    return full


@pytest.mark.parametrize("ph,pl", [(1, 0), (0, 1), (4, 2), (2, 4), (3, 0)])
def test_2q_matches_kronecker(ph, pl):
    n = 5
    m4 = haar_unitary(4, SplitMix64(31 + ph * 7 + pl))
    st = rand_state(n)
    want = two_qubit_full(n, ph, pl, m4) @ st
    oracle.apply_2q(st, ph, pl, m4)
    assert np.abs(st - want).max() < 1e-14


@pytest.mark.parametrize("p1,p2", [(0, 1), (1, 4), (2, 3)])
def test_diag_index_convention(p1, p2):
    n = 5
    d = np.exp(1j * np.array([0.3, 1.1, -0.7, 2.9]))
    st = rand_state(n)
    want = two_qubit_full(n, p1, p2, np.diag(d)) @ st
    oracle.apply_diag(st, p1, p2, d)
    assert np.abs(st - want).max() < 1e-15


@pytest.mark.parametrize("p1,p2", [(0, 1), (1, 4), (0, 3)])
def test_swaplike_acts_on_01_10(p1, p2):
    """m acts on (amplitude with bit p1 set, amplitude with bit p2 set), p1 < p2."""
    n = 5
    m = haar_unitary(2, SplitMix64(p1 * 9 + p2))
    # 4x4 in the 2*b(p2) + b(p1) basis: |01> (bit p1) and |10> (bit p2) mix through m.
    m4 = np.eye(4, dtype=complex)
    m4[1, 1], m4[1, 2], m4[2, 1], m4[2, 2] = m[0, 0], m[0, 1], m[1, 0], m[1, 1]
    st = rand_state(n)
    want = two_qubit_full(n, p2, p1, m4) @ st
    oracle.apply_swaplike(st, p1, p2, m)
    assert np.abs(st - want).max() < 1e-14


def test_swap_exact_vectors():
    """apply_swap_gate_test.hpp:65-106: swapping (0,1) and (0,2) on basis-labelled amplitudes."""
    X = np.array([[0, 1], [1, 0]], dtype=complex)
    st = np.arange(8).astype(np.complex128)
    oracle.apply_swaplike(st, 0, 1, X)
    assert list(st.real) == [0, 2, 1, 3, 4, 6, 5, 7]
    st = np.arange(8).astype(np.complex128)
    oracle.apply_swaplike(st, 0, 2, X)
    assert list(st.real) == [0, 4, 2, 6, 1, 5, 3, 7]


def test_swap_equals_three_cnots_exactly():
    """apply_swap_gate_test.hpp:110-247: SWAP == CX(a,b) CX(b,a) CX(a,b) with zero difference."""
    X = np.array([[0, 1], [1, 0]], dtype=complex)
    for (a, b) in [(0, 3), (3, 0), (1, 2)]:
        s1 = rand_state(4)
        s2 = s1.copy()
        oracle.apply_swaplike(s1, min(a, b), max(a, b), X)
        oracle.apply_c1q(s2, a, b, X)
        oracle.apply_c1q(s2, b, a, X)
        oracle.apply_c1q(s2, a, b, X)
        assert np.array_equal(s1, s2)


def test_permute_matches_tensor_transpose():
    n = 5
    st = rand_state(n)
    new_map = [2, 0, 4, 1, 3]  # program qubit k -> data position new_map[k]
    imap = np.argsort(new_map)
    out = st.copy()
    oracle.permute(out, list(range(n)), imap)
    # data index d: bit new_map[k] of d carries program bit k.
    for d in range(1 << n):
        prog = sum(((d >> new_map[k]) & 1) << k for k in range(n))
        assert out[d] == st[prog]


def test_reductions_brute_force():
    n = 6
    st = rand_state(n) * 1.3
    p = np.abs(st) ** 2
    for pos in range(n):
        assert abs(oracle.prob1(st, pos) - p[((np.arange(64) >> pos) & 1) == 1].sum()) < 1e-14
    assert abs(oracle.norm(st) - np.linalg.norm(st)) < 1e-14
    y = 0b101100
    sgn = np.array([(-1) ** bin(i & y).count("1") for i in range(64)])
    assert abs(oracle.parity_norm(st, y) - (sgn * p).sum()) < 1e-14
    b = rand_state(n)
    assert abs(oracle.overlap(st, b) - np.vdot(b, st)) < 1e-14


def run(ops, n, init="base", prec="dp", graph=None):
    c = Circuit(n, precision=prec, init=init, graph=graph or [])
    for op in ops:
        c.add(*op)
    return oracle.run(c.to_text())


def test_reference_closed_forms_1q():
    """apply_1q_gate_test.hpp:48-180 on the last of 4 qubits (|0>: index 0, |1>: index 8)."""
    s2 = math.sqrt(2.0)
    st, _ = run([("ApplyHadamard", 3)], 4)
    assert abs(st[0] - 1 / s2) < 1e-15 and abs(st[8] - 1 / s2) < 1e-15
    c = Circuit(4, base_index=8)
    c.add("ApplyHadamard", 3)
    st, _ = oracle.run(c.to_text())
    assert abs(st[0] - 1 / s2) < 1e-15 and abs(st[8] + 1 / s2) < 1e-15
    for name, th, want0, want1 in [
        ("ApplyRotationX", 0.83, math.cos(0.83 / 2), -1j * math.sin(0.83 / 2)),
        ("ApplyRotationY", 0.75, math.cos(0.75 / 2), math.sin(0.75 / 2)),
        ("ApplyRotationZ", 0.35, complex(math.cos(0.35 / 2), -math.sin(0.35 / 2)), 0),
    ]:
        st, _ = run([(name, 3, th)], 4)
        assert abs(st[0] - want0) < 1e-15 and abs(st[8] - want1) < 1e-15


def test_reference_expectations_and_probabilities():
    """expectation_values_test.hpp / measure_test.hpp: +-1, 0 and 0.5 outcomes."""
    c = Circuit(3)
    c.add("ApplyPauliX", 1)
    c.add("ApplyHadamard", 2)
    c.add("GetProbability", 0)
    c.add("GetProbability", 1)
    c.add("GetProbability", 2)
    c.add("ExpectationValue", 1.0, 1, [1, 3])
    c.add("ExpectationValue", 1.0, 1, [2, 1])
    c.add("ExpectationValue", 1.0, 1, [0, 2])
    c.add("ExpectationValue", 1.0, 2, [1, 3, 2, 1])
    _, res = oracle.run(c.to_text())
    vals = [float(l.split()[-1]) for l in res]
    assert vals[:3] == pytest.approx([0.0, 1.0, 0.5], abs=1e-15)
    assert vals[3:] == pytest.approx([-1.0, 1.0, 0.0, -1.0], abs=1e-14)


def test_qft_equals_dft():
    """examples/quantum_fourier_transform.cpp:28-73: H + controlled phases + swaps == DFT."""
    n = 5
    st0 = rand_state(n)
    c = Circuit(n)
    for q in reversed(range(n)):
        c.add("ApplyHadamard", q)
        for k in reversed(range(q)):
            c.add("ApplyCPhaseRotation", k, q, math.pi / (1 << (q - k)))
    for q in range(n // 2):
        c.add("ApplySwap", q, n - 1 - q)
    # Feed st0 in via amplitudes: run the gates on each basis state and combine (linearity).
    N = 1 << n
    U = np.zeros((N, N), dtype=complex)
    for j in range(N):
        c.base_index = j
        U[:, j], _ = oracle.run(c.to_text())
    want = np.fft.ifft(st0) * math.sqrt(N)
    assert np.abs(U @ st0 - want).max() < 1e-13


def test_maxcut_diag_and_qaoa_closed_forms():
    """qaoa_features_test.hpp:39-150: 6-ring cut values; weighted cut 2.4 at index 4, 4.8 at 36."""
    ring = [(i, (i + 1) % 6, 1.0) for i in range(6)]
    c = Circuit(6, graph=ring)
    c.add("QaoaPhase", 0.0)
    for k in (0, 63, 4, 36, 21):
        c.add("GetGlobalAmplitude", k)
    c.add("QaoaExpectation")
    _, res = oracle.run(c.to_text())
    # |0> has cut 0; QaoaExpectation of |0> is 0.
    assert float(res[-1].split()[-1]) == 0.0
    wring = [(0, 1, 1.0), (1, 2, 1.0), (2, 3, 1.4), (3, 4, 1.0), (4, 5, 1.0), (5, 0, 1.4)]
    for idx, cut in [(4, 2.4), (36, 4.8), (0, 0.0), (63, 0.0)]:
        c = Circuit(6, base_index=idx, graph=wring)
        c.add("QaoaExpectation")
        _, res = oracle.run(c.to_text())
        assert abs(float(res[-1].split()[-1]) - cut) < 1e-15
    for idx in range(64):  # ring: every cut is even and no single edge can be cut
        c = Circuit(6, base_index=idx, graph=ring)
        c.add("QaoaExpectation")
        _, res = oracle.run(c.to_text())
        v = float(res[-1].split()[-1])
        bits = [(idx >> i) & 1 for i in range(6)]
        assert v == sum(bits[i] != bits[(i + 1) % 6] for i in range(6))


def test_toffoli_truth_table():
    for j in range(8):
        c = Circuit(3, base_index=j)
        c.add("ApplyToffoli", 0, 1, 2)
        st, _ = oracle.run(c.to_text())
        want = j ^ 4 if (j & 3) == 3 else j
        assert abs(st[want] - 1) < 1e-15 and np.sum(np.abs(st) > 1e-15) == 1


def test_unitaries_preserve_norm_sp_and_dp():
    for name in ("brick10_dp", "brick10_sp"):
        import os
        text = open(os.path.join(os.path.dirname(__file__), "golden", name + ".circ")).read()
        st, _ = oracle.run(text)
        tol = 1e-12 if st.dtype == np.complex128 else 1e-5
        assert abs(np.linalg.norm(st.astype(complex)) - 1) < tol
