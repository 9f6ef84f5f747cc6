"""State statistics on the GPU: Entropy and GoogleStats (src/qureg_utils.cpp:306-450) against
the compiled reference (tests/golden/stats_*: same circuits, reference read-outs) and closed
forms; the QAOA cost histograms (src/qaoa_features.cpp:335-529) against numpy on the same
state with cut values recomputed from the edge list, plus the reference's own unit-test
identities (unit_test/include/qaoa_features_test.hpp:80-86, 154-171)."""
import glob
import math
import os

import numpy as np
import pytest

from iqs_amd import runner
from iqs_amd.circuits import Circuit

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _m():
    import intelqs_py
    return intelqs_py


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "stats*.statcirc"))), ids=os.path.basename)
def test_entropy_and_google_stats_match_reference(path):
    c = Circuit.parse(open(path).read())
    _, out = runner.run(c)
    ref = [l for l in open(path.replace(".statcirc", ".statres")).read().splitlines() if l]
    assert len(out) == len(ref)
    # complex64: the state itself agrees with the reference's to FP32 rounding (~1e-7
    # relative); a moment of order k scales that relative error by about 2k.
    rtol = 1e-12 if c.precision == "dp" else 1e-4
    for a, b in zip(out, ref):
        ta, tb = a.split(), b.split()
        assert ta[:2] == tb[:2]
        va, vb = np.array([float(x) for x in ta[2:]]), np.array([float(x) for x in tb[2:]])
        assert np.all(np.abs(va - vb) <= rtol * np.abs(vb) + 1e-300), (a, b)


@pytest.mark.parametrize("cls", ["QubitRegister", "QubitRegisterSP"])
def test_google_stats_equal_numpy_on_the_same_state(cls):
    """The reduction itself, on the register's own amplitudes (FP64 accumulation)."""
    m = _m()
    n = 11
    psi = getattr(m, cls)(n, "++++", 0, 0)
    rs = np.random.default_rng(3)
    for q in range(n):
        psi.ApplyRotationY(q, float(rs.uniform(0, 3)))
        psi.ApplyCPauliX(q, (q + 3) % n)
    p = np.abs(np.array(psi, copy=True).astype(np.complex128)) ** 2
    nz = p[p > 0]
    want = [-np.sum(nz * np.log(nz)) / math.log(2), -np.sum(np.log(nz)) / math.log(2) / 2 ** n]
    want += [np.sum(p ** k) * 2.0 ** (n * (k - 1)) / math.factorial(k) for k in range(2, 11)]
    got = psi.GoogleStats()
    assert np.all(np.abs(np.array(got) - want) <= 1e-11 * np.abs(want))
    assert abs(psi.Entropy() - want[0]) <= 1e-12 * want[0]


def test_google_stats_closed_forms():
    """|psi> = (|0> + |8>)/sqrt(2) + ... : a state with probabilities 1/2, 1/4, 1/4 has
    entropy 1.5 bits, average self-information (1 + 2 + 2) / 2^n bits and moments
    sum p^k = 2^-k + 2 * 4^-k scaled by 2^(n (k - 1)) / k!."""
    m = _m()
    n = 10
    psi = m.QubitRegister(n, "base", 0, 0)
    psi.ApplyHadamard(3)
    psi.ApplyCHadamard(3, 7)
    assert abs(psi.Entropy() - 1.5) < 1e-14
    st = psi.GoogleStats()
    assert abs(st[0] - 1.5) < 1e-14
    assert abs(st[1] - 5.0 / 2 ** n) < 1e-16
    for k in range(2, 11):
        want = (2.0 ** -k + 2 * 4.0 ** -k) * 2.0 ** (n * (k - 1)) / math.factorial(k)
        assert abs(st[k] - want) <= 1e-13 * want
    u = m.QubitRegister(n, "++++", 0, 0)
    assert abs(u.Entropy() - n) < 1e-12  # uniform distribution: n bits


def _cut_values(n, edges):
    x = np.arange(1 << n)
    c = np.zeros(1 << n)
    for u, v, w in edges:
        c += w * (((x >> u) ^ (x >> v)) & 1)
    return c


def _adjacency(n, edges):
    a = [0.0] * (n * n)
    for u, v, w in edges:
        a[u * n + v] = a[v * n + u] = w
    return a


def _qaoa_state(m, n, diag, seed):
    rs = np.random.default_rng(seed)
    psi = m.QubitRegister(n, "++++", 0, 0)
    for _ in range(3):
        m.ImplementQaoaLayerBasedOnCostFunction(psi, diag, float(rs.uniform(0, math.pi)))
        b = float(rs.uniform(0, math.pi))
        for q in range(n):
            psi.ApplyRotationX(q, 2 * b)
    return psi


def test_cost_histogram_unweighted():
    m = _m()
    n = 10
    edges = [(i, (i + 1) % n, 1.0) for i in range(n)] + [(0, 5, 1.0), (2, 7, 1.0)]
    diag = m.QubitRegister(n, "base", 0, 0)
    mx = m.InitializeVectorAsMaxCutCostFunction(diag, [int(x) for x in _adjacency(n, edges)])
    psi = _qaoa_state(m, n, diag, 5)
    h = np.array(m.GetHistogramFromCostFunction(psi, diag, mx))
    assert h.shape == (mx + 1,)
    p = np.abs(np.array(psi, copy=True)) ** 2
    cuts = _cut_values(n, edges).astype(int)
    want = np.bincount(cuts, weights=p, minlength=mx + 1)
    assert np.max(np.abs(h - want)) < 1e-14
    e = m.GetExpectationValueFromCostFunction(psi, diag)
    assert abs(e - sum(j * h[j] for j in range(len(h)))) < 1e-12
    with pytest.raises(RuntimeError):
        m.GetHistogramFromCostFunction(psi, diag, mx - 1)  # cut values above max_value


def test_cost_histogram_weighted_rounded_and_binned():
    m = _m()
    n = 9
    rs = np.random.default_rng(8)
    edges = [(i, (i + 1) % n, float(rs.uniform(0.2, 1.7))) for i in range(n)] + [(1, 6, 0.85)]
    diag = m.QubitRegister(n, "base", 0, 0)
    mx = m.InitializeVectorAsWeightedMaxCutCostFunction(diag, _adjacency(n, edges))
    psi = _qaoa_state(m, n, diag, 9)
    p = np.abs(np.array(psi, copy=True)) ** 2
    cuts = np.array(diag, copy=True).real
    assert np.max(np.abs(cuts - _cut_values(n, edges))) < 1e-12
    e = m.GetExpectationValueFromCostFunction(psi, diag)
    hr = np.array(m.GetHistogramFromCostFunctionWithWeightsRounded(psi, diag, mx))
    assert hr.shape == (int(math.floor(mx)) + 1,)
    want = np.bincount(np.floor(cuts + 1e-7).astype(int), weights=p, minlength=hr.size)
    assert np.max(np.abs(hr - want)) < 1e-14
    assert abs(e - sum(j * hr[j] for j in range(hr.size))) <= 1.0 + 1e-7
    bw = 0.1
    hb = np.array(m.GetHistogramFromCostFunctionWithWeightsBinned(psi, diag, mx, bw))
    assert hb.shape == (int(math.ceil(mx / bw)) + 1,)
    want = np.bincount(np.floor(cuts / bw + 1e-7).astype(int), weights=p, minlength=hb.size)
    assert np.max(np.abs(hb - want)) < 1e-14
    assert abs(e - sum(j * bw * hb[j] for j in range(hb.size))) <= bw + 1e-7
