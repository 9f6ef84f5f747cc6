"""Closed form of depth-1 QAOA MaxCut and its gradient (a size-independent known answer).

For |psi> = exp(-i beta sum_q X_q) exp(-i gamma C) |+>^n with C = sum_e (1 - Z_u Z_v)/2,
each edge (u, v) contributes (Wang, Hadfield, Jiang, Rieffel, Phys. Rev. A 97, 022304
(2018), Theorem 1)

  <C_uv> = 1/2 + 1/4 sin(4 beta) sin(gamma) (cos^du gamma + cos^dv gamma)
               - 1/4 sin^2(2 beta) cos^(du + dv - 2 lam) gamma (1 - cos^lam (2 gamma))

with du = deg(u) - 1, dv = deg(v) - 1 and lam the number of triangles through (u, v).
Unit weights.  The mixer exp(-i beta X) is the reference's RotationX(2 beta)
(src/qureg_apply1qubitgate.cpp:277-285).  tests/test_factored_cpu.py checks the formula
and its derivatives against brute-force state vectors.

Test infrastructure only.
"""
from __future__ import annotations

import math
from typing import Sequence, Tuple


def _edge_terms(n: int, edges: Sequence[Tuple]):
    adj = [set() for _ in range(n)]
    for e in edges:
        adj[e[0]].add(e[1])
        adj[e[1]].add(e[0])
    for e in edges:
        u, v = e[0], e[1]
        yield len(adj[u]) - 1, len(adj[v]) - 1, len(adj[u] & adj[v])


def cost(n: int, edges, gamma: float, beta: float) -> float:
    c, c2 = math.cos(gamma), math.cos(2 * gamma)
    s4b, s2b = math.sin(4 * beta), math.sin(2 * beta)
    tot = 0.0
    for du, dv, lam in _edge_terms(n, edges):
        tot += 0.5 + 0.25 * s4b * math.sin(gamma) * (c ** du + c ** dv) \
            - 0.25 * s2b * s2b * c ** (du + dv - 2 * lam) * (1 - c2 ** lam)
    return tot


def grad(n: int, edges, gamma: float, beta: float) -> Tuple[float, float]:
    """(d cost / d gamma, d cost / d beta)."""
    c, s = math.cos(gamma), math.sin(gamma)
    c2, s2 = math.cos(2 * gamma), math.sin(2 * gamma)
    s4b, c4b, s2b = math.sin(4 * beta), math.cos(4 * beta), math.sin(2 * beta)

    def pw(x, k):  # x^k and d/dx x^k with 0^0 = 1
        return x ** k, (k * x ** (k - 1) if k else 0.0)

    dg = db = 0.0
    for du, dv, lam in _edge_terms(n, edges):
        pu, dpu = pw(c, du)
        pv, dpv = pw(c, dv)
        k = du + dv - 2 * lam
        pk, dpk = pw(c, k)
        pl, dpl = pw(c2, lam)
        a = pu + pv
        da = -s * (dpu + dpv)  # d/dgamma (cos^du + cos^dv)
        b = pk * (1 - pl)
        dbg = -s * dpk * (1 - pl) + pk * (2 * s2 * dpl)
        dg += 0.25 * s4b * (math.cos(gamma) * a + s * da) - 0.25 * s2b * s2b * dbg
        db += c4b * s * a - 0.5 * s4b * b
    return dg, db
