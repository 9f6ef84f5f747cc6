"""QAOA MaxCut on the GPU: forward cost against the CPU oracle (the reference's phase layer
and RX mixer), adjoint gradient against the GPU parameter-shift rule, against central
finite differences and against a CPU parameter-shift gradient computed entirely by the
oracle; complex64 against complex128; the torch.autograd.Function."""
import math

import numpy as np
import pytest

import oracle
from iqs_amd.circuits import Circuit, SplitMix64, random_regular_graph
from iqs_amd.qaoa import MaxCutQaoa

pytestmark = pytest.mark.gpu


def params(p, seed):
    r = SplitMix64(seed)
    return [r.uniform(0, math.pi) for _ in range(p)], [r.uniform(0, math.pi) for _ in range(p)]


def oracle_cost(n, edges, gammas, betas, shift=None):
    """<C> of the same circuit through the CPU oracle; a shifted edge phase is an ApplyDiag
    diag(1, e^{-i d w}, e^{-i d w}, 1) on (u, v), a shifted mixer an extra RX."""
    c = Circuit(n, init="plus", graph=edges)
    for l, (g, b) in enumerate(zip(gammas, betas)):
        c.add("QaoaPhase", g)
        if shift and shift[0] == "edge" and shift[1] == l:
            u, v, w = edges[shift[2]]
            ph = np.exp(-1j * shift[3] * w)
            c.add("ApplyDiag", u, v, np.array([1, ph, ph, 1]))
        for q in range(n):
            c.add("ApplyRotationX", q, 2 * b)
        if shift and shift[0] == "qubit" and shift[1] == l:
            c.add("ApplyRotationX", shift[2], 2 * shift[3])
    c.add("QaoaExpectation")
    _, res = oracle.run(c.to_text())
    return float(res[-1].split()[-1])


@pytest.fixture(scope="module")
def small():
    n, p = 10, 3
    edges = random_regular_graph(n, 3, 11)
    g, b = params(p, 5)
    return n, edges, g, b


def test_forward_cost_matches_oracle(small):
    n, edges, g, b = small
    prob = MaxCutQaoa(n, edges)
    assert abs(prob.cost(g, b) - oracle_cost(n, edges, g, b)) < 1e-12


def test_adjoint_equals_parameter_shift_and_fd(small):
    n, edges, g, b = small
    prob = MaxCutQaoa(n, edges)
    cost, dg, db = prob.adjoint(g, b)
    ps, nevals = prob.parameter_shift(g, b)
    assert nevals == len(g) * 2 * (len(edges) + n)
    for l in range(len(g)):
        assert abs(dg[l] - ps[("gamma", l)]) < 1e-10
        assert abs(db[l] - ps[("beta", l)]) < 1e-10
        assert abs(dg[l] - prob.finite_difference(g, b, "gamma", l)) < 1e-6
        assert abs(db[l] - prob.finite_difference(g, b, "beta", l)) < 1e-6
    assert abs(cost - prob.cost(g, b)) < 1e-12


def test_adjoint_equals_cpu_parameter_shift(small):
    """Gradient from the oracle's forward passes alone (no GPU) vs the GPU adjoint."""
    n, edges, g, b = small
    prob = MaxCutQaoa(n, edges)
    _, dg, db = prob.adjoint(g, b)
    for l in range(len(g)):
        s = sum(0.5 * (oracle_cost(n, edges, g, b, ("edge", l, e, math.pi / 2)) -
                       oracle_cost(n, edges, g, b, ("edge", l, e, -math.pi / 2))) for e in range(len(edges)))
        assert abs(s - dg[l]) < 1e-10
        s = sum(oracle_cost(n, edges, g, b, ("qubit", l, q, math.pi / 4)) -
                oracle_cost(n, edges, g, b, ("qubit", l, q, -math.pi / 4)) for q in range(n))
        assert abs(s - db[l]) < 1e-10


def test_weighted_graph_gradient():
    n = 8
    edges = [(u, v, 0.3 + 0.2 * k) for k, (u, v, _) in enumerate(random_regular_graph(n, 3, 4))]
    g, b = params(2, 9)
    prob = MaxCutQaoa(n, edges)
    _, dg, db = prob.adjoint(g, b)
    ps, _ = prob.parameter_shift(g, b)
    for l in range(2):
        assert abs(dg[l] - ps[("gamma", l)]) < 1e-10
        assert abs(db[l] - ps[("beta", l)]) < 1e-10


def test_complex64_against_complex128():
    n, p = 16, 4
    edges = random_regular_graph(n, 3, 21)
    g, b = params(p, 22)
    c128 = MaxCutQaoa(n, edges, "c128")
    c64 = MaxCutQaoa(n, edges, "c64", renorm_every=2)
    cost1, dg1, db1 = c128.adjoint(g, b)
    cost2, dg2, db2 = c64.adjoint(g, b)
    assert abs(cost1 - cost2) < 1e-4 * max(1.0, abs(cost1))
    assert np.abs(dg1 - dg2).max() < 1e-3 and np.abs(db1 - db2).max() < 1e-3


def test_torch_autograd_function(small):
    torch = pytest.importorskip("torch")
    from iqs_amd.qaoa import torch_function
    n, edges, g, b = small
    prob = MaxCutQaoa(n, edges)
    F = torch_function()
    gt = torch.tensor(g, dtype=torch.float64, requires_grad=True)
    bt = torch.tensor(b, dtype=torch.float64, requires_grad=True)
    loss = 3.0 * F.apply(gt, bt, prob)
    loss.backward()
    cost, dg, db = MaxCutQaoa(n, edges).adjoint(g, b)
    assert abs(loss.item() - 3.0 * cost) < 1e-12
    assert np.abs(gt.grad.numpy() - 3.0 * dg).max() < 1e-10
    assert np.abs(bt.grad.numpy() - 3.0 * db).max() < 1e-10
