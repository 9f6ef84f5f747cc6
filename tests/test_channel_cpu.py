"""Chi-matrix eigensystems of the drop-in module (intelqs_py.CM4x4 / CM16x16, host code of
include/chi_matrix.hpp) against the oracle's restatement (oracle/channels.py) and the
mathematics that fixes them:
  * eigenvalues equal numpy's eigh, ordered as Eigen's ComplexEigenSolver orders them;
  * chi = sum_k p_k sign(E_k) |E_k><E_k| with the scaled eigenvectors (what ApplyChannel
    samples), so the sampled branches average to the channel
    rho' = sum_ij chi_ij sigma_i rho sigma_j^dagger on random density matrices;
  * a diagonal chi (Pauli channel) gives unit eigenvectors in Eigen's order exactly;
  * the ideal Hadamard channel has the reference's hard-coded eigensystem
    (src/chi_matrix.cpp:210-234).
No GPU is used: the module's chi-matrix classes are host code."""
import numpy as np
import pytest

import channels as och


def _m():
    import intelqs_py
    return intelqs_py


def _random_kraus(rs, d, n):
    """n Kraus operators of a random channel (isometry from a random unitary)."""
    g = rs.normal(size=(n * d, n * d)) + 1j * rs.normal(size=(n * d, n * d))
    q, _ = np.linalg.qr(g)
    return [q[i * d:(i + 1) * d, :d] for i in range(n)]


def _solve(chi):
    m = _m()
    c = (m.CM4x4 if chi.shape[0] == 4 else m.CM16x16)(chi)
    c.SolveEigenSystem()
    return c, np.array(c.GetEigenValues()), [np.array(v) for v in c.GetEigenVectors()], \
        np.array(c.GetEigenProbabilities()), np.array(c.GetEigenCumulativeProbabilities())


@pytest.mark.parametrize("d,nk", [(2, 1), (2, 3), (4, 2), (4, 5)])
def test_eigensystem_reconstructs_channel(d, nk):
    rs = np.random.default_rng(100 * d + nk)
    chi = och.chi_from_kraus(_random_kraus(rs, d, nk))
    _, w, vecs, p, cum = _solve(chi)
    wr, _, pr, cumr = och.eigensystem(chi)
    assert np.max(np.abs(w.imag)) == 0
    assert np.max(np.abs(w.real - wr)) < 1e-12
    assert np.all(np.diff(np.abs(w.real)) >= -1e-15)
    assert np.max(np.abs(p - pr)) < 1e-12 and np.max(np.abs(cum - cumr)) < 1e-12
    assert cum[-1] == 1.0
    rec = sum(p[k] * np.sign(w[k].real) * np.outer(vecs[k], vecs[k].conj()) for k in range(len(w)))
    assert np.max(np.abs(rec - chi)) < 1e-12
    for _ in range(3):
        a = rs.normal(size=(d, d)) + 1j * rs.normal(size=(d, d))
        rho = a @ a.conj().T
        rho /= np.trace(rho)
        avg = sum(p[k] * np.sign(w[k].real) * och.branch_matrix(vecs[k]) @ rho @ och.branch_matrix(vecs[k]).conj().T
                  for k in range(len(w)))
        assert np.max(np.abs(avg - och.channel_apply(chi, rho))) < 1e-12
        # Every channel here is trace preserving: the same holds for the restatement.
        assert abs(np.trace(avg) - 1) < 1e-12


def test_indefinite_chi_keeps_signs():
    """A Hermitian but indefinite chi (a Hermiticity-preserving, non-CP map): the negative
    eigenvalues keep their sign, the probabilities use |E_k|."""
    rs = np.random.default_rng(7)
    a = rs.normal(size=(4, 4)) + 1j * rs.normal(size=(4, 4))
    chi = (a + a.conj().T) / 2
    _, w, vecs, p, _ = _solve(chi)
    assert (w.real < 0).any() and (w.real > 0).any()
    assert np.allclose(p, np.abs(w.real) / np.abs(w.real).sum(), rtol=0, atol=1e-14)
    rec = sum(p[k] * np.sign(w[k].real) * np.outer(vecs[k], vecs[k].conj()) for k in range(4))
    assert np.max(np.abs(rec - chi)) < 1e-12


def test_pauli_channel_eigensystem_is_exact():
    """Diagonal chi: no rotation, unit eigenvectors (scaled by sqrt(sum |E|) = 1) in the
    order of Eigen's selection sort, bit for bit."""
    probs = [0.7, 0.1, 0.1, 0.1]
    _, w, vecs, p, cum = _solve(np.diag(probs).astype(complex))
    order = och.eigen_order(np.array(probs))
    assert order == [1, 2, 3, 0]
    assert list(w.real) == [probs[i] for i in order]
    for k, i in enumerate(order):
        e = np.zeros(4, complex)
        e[i] = 1
        assert np.array_equal(vecs[k], e)
    assert list(p) == [probs[i] for i in order]
    d16 = np.diag(np.linspace(0.01, 0.16, 16)[::-1]).astype(complex)
    d16 /= np.trace(d16)
    _, w16, v16, _, _ = _solve(d16)
    order16 = och.eigen_order(np.diag(d16).real)
    assert list(w16.real) == [d16[i, i].real for i in order16]
    for k, i in enumerate(order16):
        assert np.argmax(np.abs(v16[k])) == i and np.abs(v16[k][i] - 1) < 1e-15


def test_ideal_hadamard_eigensystem():
    m = _m()
    chi = np.zeros((4, 4), complex)
    chi[np.ix_([1, 3], [1, 3])] = 0.5
    c = m.CM4x4(chi)
    c.EigensystemOfIdealHadamardChannel()
    h = 1 / np.sqrt(2)
    assert [complex(x) for x in c.GetEigenValues()] == [1, 0, 0, 0]
    want = [[0, h, 0, h], [0, h, 0, -h], [1, 0, 0, 0], [0, 0, 1, 0]]
    assert np.array_equal(np.array(c.GetEigenVectors()), np.array(want, complex))
    assert list(c.GetEigenCumulativeProbabilities()) == [1, 1, 1, 1]
    # The branch it always takes is H itself.
    assert np.allclose(och.branch_matrix(np.array(want[0])), np.array([[1, 1], [1, -1]]) * h, atol=1e-16)
    bad = m.CM4x4(np.eye(4, dtype=complex))
    with pytest.raises(RuntimeError):
        bad.EigensystemOfIdealHadamardChannel()


def test_chi_access_and_errors():
    m = _m()
    c = m.CM4x4()
    c[1, 2] = 0.25 + 0.5j
    assert c[1, 2] == 0.25 + 0.5j and c[0, 0] == 0
    with pytest.raises(IndexError):
        c[4, 0]
    with pytest.raises(RuntimeError):
        m.CM16x16(np.eye(4, dtype=complex))
    c[2, 1] = 0.0
    with pytest.raises(RuntimeError):  # not Hermitian
        c.SolveEigenSystem()
    assert repr(m.CM4x4()) == "<ChiMatrix for 1-qubit channel>"
    assert repr(m.CM16x16()) == "<ChiMatrix for 2-qubit channel>"


def test_oracle_restatement_pins():
    """The oracle's own pieces against closed forms: the Pauli basis is orthogonal with
    tr(s_i s_j) = d delta_ij; chi_from_kraus of a unitary U = sum c_i sigma_i reproduces
    U rho U^dagger; Eigen's ordering of [3, -1, 2, 1] is [-1, 1, 2, 3] by magnitude with
    the first minimum taken."""
    for nq, d in ((1, 2), (2, 4)):
        s = och.pauli_basis(nq)
        g = np.array([[np.trace(a.conj().T @ b) for b in s] for a in s])
        assert np.allclose(g, d * np.eye(len(s)))
    rs = np.random.default_rng(3)
    u, _ = np.linalg.qr(rs.normal(size=(4, 4)) + 1j * rs.normal(size=(4, 4)))
    a = rs.normal(size=(4, 4)) + 1j * rs.normal(size=(4, 4))
    rho = a @ a.conj().T
    assert np.allclose(och.channel_apply(och.chi_from_kraus([u]), rho), u @ rho @ u.conj().T, atol=1e-12)
    assert och.eigen_order(np.array([3.0, -1.0, 2.0, 1.0])) == [1, 3, 2, 0]
    psi = rs.normal(size=8) + 1j * rs.normal(size=8)
    x = np.array([[0, 1], [1, 0]], complex)
    assert np.allclose(och.apply_gate(psi, [2], x), psi.reshape(2, 4)[::-1].reshape(-1))
    cx = np.eye(4, dtype=complex)[[0, 1, 3, 2]]  # control = first (high) index bit
    got = och.apply_gate(psi, [0, 1], cx)
    want = psi.copy()
    for i in range(8):
        if i & 1:
            want[i] = psi[i ^ 2]
    assert np.allclose(got, want)
