"""The tile-pass schedule of the batched HEA kernel (intel-qs_amd/csrc/iqs_batch.hip),
restated on the CPU.  A layer U_l = CZ_l . prod_q RZ_q(phi_lq) . prod_q RY_q(theta_lq) is
split into real RY rotations and one diagonal per register phase: pass j, in register phases
of log2(R) consecutive local bits, applies RY_{j-1} on the bits of its locality, a diagonal
(RZ_{j-1} of those bits, CZ pairs of layer j-1, and in phase 0 the RZ_{j-1} of the qubits
the previous pass rotated: the deferred RZs), then RY_j.  Checked here:
  * every RY and every RZ is applied exactly once and every CZ pair once per layer; each RZ
    sits after its qubit's RY of the same layer and before the RY of the next; each CZ pair
    after the two RYs of its layer and before the two of the next (all sizes up to 26
    qubits, every tile shape);
  * the register ranges of every phase hold the bits they gate;
  * executing the schedule on a state vector gives the circuit's state, and the adjoint
    formulas (dC/dtheta = Im<l|Y|p> just after the RY, dC/dphi = Im<l|Z|p> at the diagonal
    holding the RZ) match central finite differences of the plain circuit.
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

    def gated(self):
        return [self.gq(b) for b in range(self.la, self.lb)]


def ops_fwd(g, pre, post, cz, deferred):
    """One pass: ("RY", layer, q) and ("D", layer, rz-qubits, cz-mask) in execution order."""
    ops = []
    for k in range(g.K):
        nb, ne, a = g.nb(k), g.ne(k), g.a(k)
        assert 0 <= a <= nb and ne <= a + g.LR <= g.M
        if pre is not None:
            ops += [("RY", pre, g.gq(b)) for b in range(nb, ne)]
            zs = [g.gq(b) for b in range(nb, ne)] + (list(deferred) if k == 0 else [])
            ops.append(("D", pre, zs, g.czmask(k) if cz else 0))
        if post is not None:
            lo, hi = max(nb - 1, g.la), (g.lb if k == g.K - 1 else ne - 1)
            assert lo >= a and hi <= a + g.LR
            ops += [("RY", post, g.gq(b)) for b in range(lo, hi)]
    return ops


def schedule(n, D, M, LR):
    """Forward gate sequence of one register as the kernel executes it."""
    if n <= M:
        g = Geo(M, LR, n, 0, 0, n, 0)
        return [op for l in range(D) for op in ops_fwd(g, l, None, True, [])]
    c = 2 * M - n
    seq, prev = [], []
    for j in range(D + 1):
        B = j & 1
        g = Geo(M, LR, n, 2 if B else 1, c if B else 0, M, c)
        seq += ops_fwd(g, j - 1 if j >= 1 else None, j if j < D else None, j >= 1, prev)
        prev = g.gated()  # this pass's RY_j leave their RZ_j to the next pass
    return seq


VARIANTS = [(13, 4), (12, 4), (12, 3), (11, 3), (10, 3), (8, 2)]  # (tile qubits, log2 amplitudes per thread)


@pytest.mark.parametrize("M,LR", VARIANTS)
def test_schedule_covers_circuit_in_order(M, LR):
    for n in range(1, min(2 * M, 26) + 1):
        for D in (1, 2, 3, 5):
            seq = schedule(n, D, M, LR)
            ry, rz = {}, {}
            for i, op in enumerate(seq):
                if op[0] == "RY":
                    assert (op[1], op[2]) not in ry
                    ry[(op[1], op[2])] = i
                else:
                    for q in op[2]:
                        assert (op[1], q) not in rz
                        rz[(op[1], q)] = i
            assert len(ry) == n * D and len(rz) == n * D
            for (l, q), i in rz.items():
                assert ry[(l, q)] < i
                if l + 1 < D:
                    assert i < ry[(l + 1, q)]
            cnt = [0] * max(n - 1, 1)
            for i, op in enumerate(seq):
                if op[0] != "D":
                    continue
                for p in range(n - 1):
                    if (op[3] >> p) & 1:
                        assert cnt[p] == op[1]
                        cnt[p] += 1
                        for q in (p, p + 1):
                            assert ry[(op[1], q)] < i
                            if op[1] + 1 < D:
                                assert i < ry[(op[1] + 1, q)]
            assert all(x == D for x in cnt[: n - 1])


def ry_mat(th):
    c, s = math.cos(th / 2), math.sin(th / 2)
    return np.array([[c, -s], [s, c]], dtype=complex)


def u_mat(th, ph):
    e = np.exp(-0.5j * ph)
    return np.diag([e, np.conj(e)]) @ ry_mat(th)


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


def diag_of(n, zs, czm, ph_row):
    """The diagonal of op ("D", l, zs, czm): RZ(phi_lq) on q in zs, then the CZ pairs."""
    x = np.arange(1 << n)
    d = czd(n, czm).astype(complex)
    for q in zs:
        d = d * np.exp(0.5j * ph_row[q] * (2 * ((x >> q) & 1) - 1))
    return d


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
        if op[0] == "RY":
            psi = apply1(psi, op[2], ry_mat(th[op[1], op[2]]))
        else:
            psi = psi * diag_of(n, op[2], op[3], ph[op[1]])
    o = n - 2 * np.array([bin(v).count("1") for v in range(1 << n)])
    cost = float(np.sum(o * np.abs(psi) ** 2))
    lam = o * psi
    g = np.zeros(th.shape + (2,))
    Y = np.array([[0, -1j], [1j, 0]])
    x = np.arange(1 << n)
    for op in reversed(seq):
        if op[0] == "RY":
            l, q = op[1], op[2]
            g[l, q, 0] = np.imag(np.vdot(lam, apply1(psi, q, Y)))
            rd = ry_mat(th[l, q]).T  # RY^dagger (real)
            psi, lam = apply1(psi, q, rd), apply1(lam, q, rd)
        else:
            t = np.imag(np.conj(lam) * psi)
            for q in op[2]:
                g[op[1], q, 1] = float(np.sum(t * (1 - 2 * ((x >> q) & 1))))  # Im<l|Z_q|p>
            d = np.conj(diag_of(n, op[2], op[3], ph[op[1]]))
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
