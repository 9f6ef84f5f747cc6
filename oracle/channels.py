"""numpy restatement of the reference's quantum channels (src/chi_matrix.cpp:66-205,
src/qureg_apply_channel.cpp:17-106).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  The product never imports it.

The reference diagonalizes chi with Eigen's ComplexEigenSolver (a third-party dependency
that is absent here, and without which the reference's noise build does not compile).
Restated here: chi is Hermitian, so its eigenpairs are numpy's `eigh`; Eigen orders them by
ascending |eigenvalue| with a selection that swaps the first minimum of the remaining tail
into place (ComplexEigenSolver::sortEigenvalues), which `eigen_order` restates.  Branch
matrices are written from their definition, sum_i e_i sigma_i, not from the reference's
expanded entries, so the expansion in the product is checked against the definition.
"""
from __future__ import annotations

import itertools

import numpy as np

PAULI = [
    np.eye(2, dtype=complex),
    np.array([[0, 1], [1, 0]], dtype=complex),
    np.array([[0, -1j], [1j, 0]], dtype=complex),
    np.array([[1, 0], [0, -1]], dtype=complex),
]


def pauli_basis(nq: int):
    """sigma_i for i in [0, 4^nq): for two qubits sigma_{4a+b} = sigma_a (x) sigma_b, the
    first factor on the high bit of the 4x4 index 2 b(qubit1) + b(qubit2)
    (qureg_apply_channel.cpp:78-86, chi_matrix.hpp:17-21)."""
    out = []
    for idx in itertools.product(range(4), repeat=nq):
        m = np.array([[1.0 + 0j]])
        for a in idx:
            m = np.kron(m, PAULI[a])
        out.append(m)
    return out


def channel_apply(chi: np.ndarray, rho: np.ndarray) -> np.ndarray:
    """rho' = sum_ij chi_ij sigma_i rho sigma_j^dagger (chi_matrix.hpp:14)."""
    nq = 1 if chi.shape[0] == 4 else 2
    s = pauli_basis(nq)
    out = np.zeros_like(rho)
    for i, j in itertools.product(range(len(s)), repeat=2):
        if chi[i, j] != 0:
            out += chi[i, j] * s[i] @ rho @ s[j].conj().T
    return out


def chi_from_kraus(kraus) -> np.ndarray:
    """chi of the channel rho -> sum_K K rho K^dagger: K = sum_i c_i sigma_i gives
    chi_ij = sum_K c_i conj(c_j), with c_i = tr(sigma_i K) / d."""
    d = kraus[0].shape[0]
    s = pauli_basis(1 if d == 2 else 2)
    chi = np.zeros((len(s), len(s)), dtype=complex)
    for K in kraus:
        c = np.array([np.trace(p.conj().T @ K) / d for p in s])
        chi += np.outer(c, c.conj())
    return chi


def eigen_order(w: np.ndarray) -> list:
    """Permutation that Eigen's sortEigenvalues applies: for i = 0.., the first index of the
    smallest |w| among positions i.. is swapped with position i."""
    idx = list(range(len(w)))
    a = [abs(x) for x in w]
    for i in range(len(w)):
        k = i + int(np.argmin(a[i:]))
        if k != i:
            a[i], a[k] = a[k], a[i]
            idx[i], idx[k] = idx[k], idx[i]
    return idx


def eigensystem(chi: np.ndarray):
    """(eigenvalues, scaled eigenvectors, probabilities, cumulative probabilities), as
    SolveEigenSystem + NormalizeEigenProbAndRenormalizeEigenVect leave them
    (chi_matrix.cpp:66-127, 166-205), eigenvectors with the phase of numpy's eigh."""
    w, v = np.linalg.eigh(chi)
    order = eigen_order(w)
    w, vecs = w[order], [v[:, k] for k in order]
    p = np.abs(w)
    norm = p.sum()
    cum = np.cumsum(p)
    if norm != 0 and norm != 1:
        p, cum = p / norm, cum / norm
        vecs = [x * np.sqrt(norm) for x in vecs]
    return w, vecs, p, cum


def branch(cum, r: float) -> int:
    """First k with r <= cum[k] (qureg_apply_channel.cpp:23-31)."""
    k = 0
    while r > cum[k]:
        k += 1
    return k


def branch_matrix(e) -> np.ndarray:
    """The matrix sum_i e_i sigma_i of one scaled eigenvector (2x2 or 4x4)."""
    s = pauli_basis(1 if len(e) == 4 else 2)
    return sum(c * p for c, p in zip(e, s))


def apply_gate(psi: np.ndarray, qubits, m: np.ndarray) -> np.ndarray:
    """m on `qubits` (first = high bit of m's index) of a state vector, qubit q = bit q."""
    n = int(np.log2(psi.size))
    t = psi.reshape([2] * n)
    axes = [n - 1 - q for q in qubits]
    k = len(qubits)
    t = np.moveaxis(t, axes, list(range(k)))
    sh = t.shape
    t = (m @ t.reshape(2 ** k, -1)).reshape(sh)
    return np.moveaxis(t, list(range(k)), axes).reshape(-1)
