"""ApplyChannel on the GPU register (src/qureg_apply_channel.cpp of the reference) against the
oracle's restatement (oracle/channels.py): for each seed, the branch k is the one the
uniform draw of the register's state stream selects from the cumulative probabilities, and
the state after the call must be sum_i e_i sigma_i |psi> for that branch's eigenvector --
to 1e-12 (complex128) / 1e-5 (complex64) after fixing the global phase that the eigenvector
carries (an eigenvector is defined up to a phase).  Pauli channels (diagonal chi) have no
such freedom and are compared without it.  The overall sign of channels flips exactly on
the branches with a negative eigenvalue."""
import numpy as np
import pytest

import channels as och

pytestmark = pytest.mark.gpu

N = 6


def _m():
    import intelqs_py
    return intelqs_py


def _random_state(rs, n):
    v = rs.normal(size=2 ** n) + 1j * rs.normal(size=2 ** n)
    return v / np.linalg.norm(v)


def _register(psi, sp=False):
    m = _m()
    reg = (m.QubitRegisterSP if sp else m.QubitRegister)(N, "base", 0, 0)
    view = np.array(reg, copy=False)
    view[:] = psi.astype(view.dtype)
    return reg


def _draw(seed):
    m = _m()
    r = m.RandomNumberGenerator()
    r.SetSeedStreamPtrs(seed)
    return r.GetUniformRandomNumbers(1, 0.0, 1.0, "state")[0]


def _apply(reg, seed, qubits, chi_obj):
    m = _m()
    rng = m.RandomNumberGenerator()
    rng.SetSeedStreamPtrs(seed)
    reg.SetRngPtr(rng)
    reg.ApplyChannel(*qubits, chi_obj)
    reg.ResetRngPtr()
    return np.array(reg, copy=True).astype(np.complex128)


def _phase_fixed(got, want):
    ov = np.vdot(want, got)
    return got * (np.conj(ov) / abs(ov)) if abs(ov) > 0 else got


def _kraus_chi(rs, d, nk):
    g = rs.normal(size=(nk * d, nk * d)) + 1j * rs.normal(size=(nk * d, nk * d))
    q, _ = np.linalg.qr(g)
    return och.chi_from_kraus([q[i * d:(i + 1) * d, :d] for i in range(nk)])


@pytest.mark.parametrize("qubits,d,nk", [((2,), 2, 3), ((0,), 2, 2), ((5, 1), 4, 3), ((0, 4), 4, 2)])
def test_channel_branch_states(qubits, d, nk):
    rs = np.random.default_rng(sum(qubits) * 10 + d + nk)
    chi = _kraus_chi(rs, d, nk)
    m = _m()
    c = (m.CM4x4 if d == 2 else m.CM16x16)(chi)
    c.SolveEigenSystem()
    vecs = [np.array(v) for v in c.GetEigenVectors()]
    cum = np.array(c.GetEigenCumulativeProbabilities())
    psi = _random_state(rs, N)
    seen = set()
    for seed in range(1, 25):
        k = och.branch(cum, _draw(seed))
        seen.add(k)
        want = och.apply_gate(psi, list(qubits), och.branch_matrix(vecs[k]))
        got = _apply(_register(psi), seed, qubits, c)
        assert np.max(np.abs(got - want)) < 1e-12, (seed, k)
    assert len(seen) >= 2
    # The oracle's own eigensystem selects branches with the same projectors.
    wr, vr, _, cumr = och.eigensystem(chi)
    for seed in range(1, 9):
        k = och.branch(cumr, _draw(seed))
        want = och.apply_gate(psi, list(qubits), och.branch_matrix(vr[k]))
        got = _apply(_register(psi), seed, qubits, c)
        assert np.max(np.abs(_phase_fixed(got, want) - want)) < 1e-12, (seed, k)


def test_pauli_channel_matches_without_phase_freedom():
    rs = np.random.default_rng(11)
    probs = np.array([0.55, 0.2, 0.15, 0.1])
    m = _m()
    c = m.CM4x4(np.diag(probs).astype(complex))
    c.SolveEigenSystem()
    wr, vr, _, cumr = och.eigensystem(np.diag(probs).astype(complex))
    psi = _random_state(rs, N)
    for seed in range(1, 20):
        k = och.branch(cumr, _draw(seed))
        want = och.apply_gate(psi, [3], och.branch_matrix(vr[k]))
        got = _apply(_register(psi), seed, (3,), c)
        assert np.max(np.abs(got - want)) < 1e-12, (seed, k)


def test_hadamard_channel_is_hadamard():
    m = _m()
    chi = np.zeros((4, 4), complex)
    chi[np.ix_([1, 3], [1, 3])] = 0.5
    c = m.CM4x4(chi)
    c.EigensystemOfIdealHadamardChannel()
    rs = np.random.default_rng(5)
    psi = _random_state(rs, N)
    h = np.array([[1, 1], [1, -1]]) / np.sqrt(2)
    for seed in (1, 2, 3):
        got = _apply(_register(psi), seed, (4,), c)
        assert np.max(np.abs(got - och.apply_gate(psi, [4], h))) < 1e-12


def test_sign_of_channels_and_complex64():
    rs = np.random.default_rng(9)
    a = rs.normal(size=(4, 4)) + 1j * rs.normal(size=(4, 4))
    chi = (a + a.conj().T) / 8
    m = _m()
    c = m.CM4x4(chi)
    c.SolveEigenSystem()
    w = np.array(c.GetEigenValues()).real
    cum = np.array(c.GetEigenCumulativeProbabilities())
    vecs = [np.array(v) for v in c.GetEigenVectors()]
    assert (w < 0).any()
    psi = _random_state(rs, N)
    for sp in (False, True):
        reg = _register(psi, sp)
        gen = m.RandomNumberGeneratorSP if sp else m.RandomNumberGenerator
        rng = gen()
        rng.SetSeedStreamPtrs(4242)
        reg.SetRngPtr(rng)
        ref = gen()
        ref.SetSeedStreamPtrs(4242)
        sign, state = 1.0, psi.copy()
        for step in range(12):
            k = och.branch(cum, ref.GetUniformRandomNumbers(1, 0.0, 1.0, "state")[0])
            sign *= -1.0 if w[k] < 0 else 1.0
            state = och.apply_gate(state, [step % N], och.branch_matrix(vecs[k]))
            reg.ApplyChannel(step % N, c)
            assert reg.GetOverallSignOfChannels() == sign
        reg.ResetRngPtr()
        got = np.array(reg, copy=True).astype(np.complex128)
        tol = 1e-5 * np.max(np.abs(state)) * 10 if sp else 1e-11 * np.max(np.abs(state))
        assert np.max(np.abs(got - state)) < tol


def test_channel_errors():
    m = _m()
    reg = m.QubitRegister(3, "base", 0, 0)
    c = m.CM4x4(np.eye(4, dtype=complex) / 4)
    with pytest.raises(RuntimeError):
        reg.ApplyChannel(0, c)  # no generator attached
    rng = m.RandomNumberGenerator()
    rng.SetSeedStreamPtrs(1)
    reg.SetRngPtr(rng)
    with pytest.raises(RuntimeError):
        reg.ApplyChannel(0, c)  # eigensystem not computed
    z = m.CM4x4()
    z.SolveEigenSystem()
    with pytest.raises(RuntimeError):
        reg.ApplyChannel(0, z)  # zero chi-matrix
    reg.ResetRngPtr()
