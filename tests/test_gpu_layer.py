"""Layers of one gate on every qubit (iqsh_apply_1q_all / iqsh_adjoint_1q_all): the forward
layer against the oracle applying the gate qubit by qubit (20 qubits and up: register
passes of one tile per workgroup; below: the grid-stride form); the adjoint layer against the
per-qubit adjoint step on the GPU and against sum_q 2 Im <lam| G_q |psi> computed with the
oracle's 1q kernel on the host (G_q psi for each q), then the oracle's per-qubit undo."""
import numpy as np
import pytest

import oracle
from iqs_amd.circuits import SplitMix64
from iqs_amd._lib import IqsError
from iqs_amd.engine import StateVector

pytestmark = pytest.mark.gpu
TOL = {"c128": 1e-12, "c64": 1e-5}
DT = {"c128": np.complex128, "c64": np.complex64}


def _state(n, dt, seed):
    r = np.random.default_rng(seed)
    v = r.normal(size=1 << n) + 1j * r.normal(size=1 << n)
    return (v / np.linalg.norm(v)).astype(DT[dt])


def _unitary(seed):
    r = SplitMix64(seed)
    a, b, c = (r.uniform(0, 6.28) for _ in range(3))
    return np.array([[np.exp(1j * a) * np.cos(b), -np.exp(1j * c) * np.sin(b)],
                     [np.exp(-1j * c) * np.sin(b), np.exp(-1j * a) * np.cos(b)]])


def _gpu(h, dt):
    sv = StateVector(int(h.size).bit_length() - 1, dt)
    sv.upload(h)
    return sv


@pytest.mark.parametrize("dt", ["c128", "c64"])
@pytest.mark.parametrize("n", [10, 11, 13, 17, 20])
def test_layer_forward(dt, n):
    h = _state(n, dt, n)
    u = _unitary(100 + n)
    sv = _gpu(h, dt)
    sv.apply_1q_all(u)
    for q in range(n):
        oracle.apply_1q(h, q, u)
    assert np.abs(sv.download() - h).max() < TOL[dt]


@pytest.mark.parametrize("gen", ["general", "X"])
@pytest.mark.parametrize("dt", ["c128", "c64"])
@pytest.mark.parametrize("n", [10, 12, 15, 20])
def test_layer_adjoint(dt, n, gen):
    """Any generator, and G = X exactly (the kernels' two-product form of the term)."""
    psi = _state(n, dt, 2 * n)
    lam = _state(n, dt, 2 * n + 1)
    u = _unitary(7 * n)
    uinv = u.conj().T
    if gen == "X":
        G = np.array([[0, 1], [1, 0]], dtype=complex)
    else:
        G = np.array([[0.3, 1.2 - 0.4j], [1.2 + 0.4j, -0.7]])  # any 2x2, Hermitian here
    # reference value: sum_q 2 Im <lam| G_q |psi> on the host (FP64)
    want = 0.0
    for q in range(n):
        gp = psi.astype(np.complex128).copy()
        oracle.apply_1q(gp, q, G)
        want += 2 * np.imag(np.vdot(lam.astype(np.complex128), gp))
    a, b = _gpu(psi, dt), _gpu(lam, dt)
    got = a.adjoint_1q_all(b, uinv, G)
    # per-qubit adjoint steps on the GPU
    c, d = _gpu(psi, dt), _gpu(lam, dt)
    seq = sum(c.adjoint_1q(d, q, uinv, G) for q in reversed(range(n)))
    scale = max(1.0, abs(want))
    assert abs(got - want) < (1e-11 if dt == "c128" else 2e-5) * scale
    assert abs(got - seq) < (1e-11 if dt == "c128" else 2e-5) * scale
    hp, hl = psi.copy(), lam.copy()
    for q in range(n):
        oracle.apply_1q(hp, q, uinv)
        oracle.apply_1q(hl, q, uinv)
    assert np.abs(a.download() - hp).max() < TOL[dt]
    assert np.abs(b.download() - hl).max() < TOL[dt]
    assert np.abs(a.download() - c.download()).max() < TOL[dt]


def test_layer_needs_ten_qubits():
    sv = StateVector(9, "c128")
    with pytest.raises(IqsError):
        sv.apply_1q_all(np.eye(2))


def _graph(n, seed, weighted):
    r = np.random.default_rng(seed)
    edges = [(i, (i + 1) % n, float(r.uniform(0.3, 1.9)) if weighted else 1.0) for i in range(n)]
    edges += [(0, n // 2, 0.7 if weighted else 1.0), (1, n - 2, 1.3 if weighted else 1.0)]
    return edges


@pytest.mark.parametrize("dt", ["c128", "c64"])
@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("n", [12, 20])
def test_qaoa_layer_fused_equals_unfused(dt, weighted, n):
    """iqsh_qaoa_layer = iqsh_qaoa_phase then the mixer on every qubit; the fused adjoint =
    iqsh_adjoint_1q_all then iqsh_qaoa_phase_adjoint (gradients and states); the forward
    layer also against the oracle (cost diagonal from the edge list, 1q kernel per qubit)."""
    from iqs_amd import gates as Gt
    edges = _graph(n, 3, weighted)
    gamma, beta = 0.83, 0.41
    psi = _state(n, dt, 41)
    a, b = _gpu(psi, dt), _gpu(psi, dt)
    a.qaoa_layer(edges, gamma, Gt.rx(2 * beta))
    b.qaoa_phase(edges, gamma)
    b.apply_1q_all(Gt.rx(2 * beta))
    assert np.abs(a.download() - b.download()).max() < TOL[dt]
    h = psi.astype(np.complex128).copy()
    x = np.arange(1 << n)
    cut = sum(w * (((x >> u) ^ (x >> v)) & 1) for u, v, w in edges)
    h *= np.exp(-1j * gamma * cut)
    for q in range(n):
        oracle.apply_1q(h, q, Gt.rx(2 * beta))
    assert np.abs(a.download() - h).max() < (1e-12 if dt == "c128" else 2e-5)
    lam = _state(n, dt, 42)
    X = np.array([[0, 1], [1, 0]])
    p1, l1 = _gpu(psi, dt), _gpu(lam, dt)
    dg1, db1 = p1.qaoa_layer_adjoint(l1, edges, gamma, Gt.rx(-2 * beta), X)
    p2, l2 = _gpu(psi, dt), _gpu(lam, dt)
    db2 = p2.adjoint_1q_all(l2, Gt.rx(-2 * beta), X)
    dg2 = p2.qaoa_phase_adjoint(l2, edges, gamma)
    tol = 1e-11 if dt == "c128" else 2e-5
    assert abs(db1 - db2) < tol * max(1, abs(db2)) and abs(dg1 - dg2) < tol * max(1, abs(dg2))
    assert np.abs(p1.download() - p2.download()).max() < TOL[dt]
    assert np.abs(l1.download() - l2.download()).max() < TOL[dt]
