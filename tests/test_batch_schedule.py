"""The tile-pass schedule of the batched HEA kernel (intel-qs_amd/csrc/iqs_batch.hip),
restated on the CPU: pass j gates its locality with R_{j-1}, the CZ ladder of layer j-1 and
R_j, in register phases of log2(R) consecutive local bits.  Checked here:
  * every rotation is applied exactly once and every CZ pair once per layer, after the two
    rotations of its layer and before the two of the next (all sizes up to 26 qubits);
  * the register ranges of every phase hold the bits they gate;
  * executing the schedule on a state vector gives the circuit's state, and the adjoint
    formulas (dC/dphi = Im<l|Z|p>, dC/dtheta = cos(phi) Im<l|Y|p> - sin(phi) Im<l|X|p>)
    match central finite differences of the plain circuit.
The arithmetic mirrors the kernel's Geo struct (phase bounds, local->global maps, CZ masks)."""
import itertools
import math

import numpy as np
import pytest


class Geo:
    def __init__(self, M, LR, n, kind, la, lb, c):
        self.M, self.LR, self.n, self.kind, self.la, self.lb, self.c = M, LR, n, kind, la, lb, c
        a0 = min(la, M - LR)
        self.ne0 = min(a0 + LR, lb)
        self.K = 1 + ((lb - self.ne0 + LR - 2) // (LR - 1) if lb > self.ne0 else 0)

    def ne(self, k):
        return min(self.ne0 + k * (self.LR - 1), self.lb)

    def nb(self, k):
        return self.ne(k - 1) if k else self.la

    def a(self, k):
        return min(self.nb(k) - 1 if k else self.la, self.M - self.LR)

    def gq(self, lbit):
        return lbit - self.c + self.M if (self.kind == 2 and lbit >= self.c) else lbit

    @staticmethod
    def rng(lo, hi):
        return 0 if hi <= lo else ((1 << hi) - 1) & ~((1 << lo) - 1)

    def czmask(self, k):
        g_ne = self.gq(self.ne(k) - 1) + 1
        m = self.rng(0 if k == 0 else self.gq(self.nb(k)) - 1, g_ne - 1)
        if k == self.K - 1:
            m |= self.rng(self.gq(self.lb - 1), self.n - 1)
        return m & self.rng(0, self.n - 1)


def ops_fwd(g, pre, post, cz):
    ops = []
    for k in range(g.K):
        nb, ne, a = g.nb(k), g.ne(k), g.a(k)
        assert 0 <= a <= nb and ne <= a + g.LR <= g.M
        if pre is not None:
            ops += [("R", pre, g.gq(b)) for b in range(nb, ne)]
        if cz:
            ops.append(("CZ", g.czmask(k)))
        if post is not None:
            lo, hi = max(nb - 1, g.la), (g.lb if k == g.K - 1 else ne - 1)
            assert lo >= a and hi <= a + g.LR
            ops += [("R", post, g.gq(b)) for b in range(lo, hi)]
    return ops


def schedule(n, D, M, LR):
    """Forward gate sequence of one register as the kernel executes it."""
    if n <= M:
        g = Geo(M, LR, n, 0, 0, n, 0)
        return [op for l in range(D) for op in ops_fwd(g, l, None, True)]
    c = 2 * M - n
    seq = []
    for j in range(D + 1):
        B = j & 1
        g = Geo(M, LR, n, 2 if B else 1, c if B else 0, M, c)
        seq += ops_fwd(g, j - 1 if j >= 1 else None, j if j < D else None, j >= 1)
    return seq


VARIANTS = [(13, 4), (12, 4), (12, 3), (11, 3), (10, 3), (8, 2)]  # (tile qubits, log2 amplitudes per thread)


@pytest.mark.parametrize("M,LR", VARIANTS)
def test_schedule_covers_circuit_in_order(M, LR):
    for n in range(1, min(2 * M, 26) + 1):
        for D in (1, 2, 3, 5):
            seq = schedule(n, D, M, LR)
            pos = {}
            for i, op in enumerate(seq):
                if op[0] == "R":
                    assert (op[1], op[2]) not in pos
                    pos[(op[1], op[2])] = i
            assert len(pos) == n * D
            cnt = [0] * max(n - 1, 1)
            for i, op in enumerate(seq):
                if op[0] != "CZ":
                    continue
                for p in range(n - 1):
                    if (op[1] >> p) & 1:
                        l = cnt[p]
                        cnt[p] += 1
                        for q in (p, p + 1):
                            assert pos[(l, q)] < i
                            if l + 1 < D:
                                assert i < pos[(l + 1, q)]
            assert all(x == D for x in cnt[: n - 1])


def u_mat(th, ph):
    c, s = math.cos(th / 2), math.sin(th / 2)
    e = np.exp(-0.5j * ph)
    return np.array([[e * c, -e * s], [np.conj(e) * s, np.conj(e) * c]])


def apply1(psi, q, m):
    x = np.arange(psi.size)
    i0 = x[((x >> q) & 1) == 0]
    a, b = psi[i0].copy(), psi[i0 | (1 << q)].copy()
    out = psi.copy()
    out[i0], out[i0 | (1 << q)] = m[0, 0] * a + m[0, 1] * b, m[1, 0] * a + m[1, 1] * b
    return out


def czd(n, mask):
    x = np.arange(1 << n)
    par = np.array([bin(v).count("1") & 1 for v in (x & (x >> 1) & mask)])
    return np.where(par == 1, -1.0, 1.0)


def plain_cost(n, th, ph):
    psi = np.zeros(1 << n, complex)
    psi[0] = 1
    for l in range(th.shape[0]):
        for q in range(n):
            psi = apply1(psi, q, u_mat(th[l, q], ph[l, q]))
        psi = psi * czd(n, (1 << (n - 1)) - 1)
    o = n - 2 * np.array([bin(v).count("1") for v in range(1 << n)])
    return float(np.sum(o * np.abs(psi) ** 2))


def run_schedule(n, seq, th, ph):
    psi = np.zeros(1 << n, complex)
    psi[0] = 1
    for op in seq:
        psi = apply1(psi, op[2], u_mat(th[op[1], op[2]], ph[op[1], op[2]])) if op[0] == "R" else psi * czd(n, op[1])
    o = n - 2 * np.array([bin(v).count("1") for v in range(1 << n)])
    cost = float(np.sum(o * np.abs(psi) ** 2))
    lam = o * psi
    g = np.zeros(th.shape + (2,))
    X = np.array([[0, 1], [1, 0]])
    Y = np.array([[0, -1j], [1j, 0]])
    Z = np.diag([1, -1])
    for op in reversed(seq):
        if op[0] == "R":
            l, q = op[1], op[2]
            sx, sy, sz = (np.imag(np.vdot(lam, apply1(psi, q, P))) for P in (X, Y, Z))
            g[l, q] = (math.cos(ph[l, q]) * sy - math.sin(ph[l, q]) * sx, sz)
            ud = u_mat(th[l, q], ph[l, q]).conj().T
            psi, lam = apply1(psi, q, ud), apply1(lam, q, ud)
        else:
            d = czd(n, op[1])
            psi, lam = psi * d, lam * d
    return cost, g


@pytest.mark.parametrize("M,LR", [(4, 2), (5, 2), (6, 3), (8, 3)])
def test_schedule_state_and_adjoint_gradient(M, LR):
    rs = np.random.default_rng(M * 10 + LR)
    for n in sorted({2, M - 1, M, M + 1, 2 * M - 1, 2 * M}):
        if n < 1 or n > 9:
            continue
        D = int(rs.integers(1, 4))
        th, ph = rs.uniform(0, 2 * math.pi, (D, n)), rs.uniform(0, 2 * math.pi, (D, n))
        cost, g = run_schedule(n, schedule(n, D, M, LR), th, ph)
        assert abs(cost - plain_cost(n, th, ph)) < 1e-12
        h = 1e-6
        for l, q, k in itertools.product(range(D), range(n), range(2)):
            a, b = th.copy(), ph.copy()
            (a if k == 0 else b)[l, q] += h
            cp = plain_cost(n, a, b)
            a, b = th.copy(), ph.copy()
            (a if k == 0 else b)[l, q] -= h
            cm = plain_cost(n, a, b)
            assert abs(g[l, q, k] - (cp - cm) / (2 * h)) < 1e-6
