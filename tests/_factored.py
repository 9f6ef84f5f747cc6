"""Closed-form checker for full-size registers: a state kept as a product of small clusters.

A circuit that starts from |0...0> and only ever entangles a few positions at a time has a
state |psi> = (x)_c |phi_c>, one small dense vector per cluster c of data positions.  This
class tracks that factorization: every gate is applied by the CPU oracle's kernels to the
cluster vector (merging the clusters a gate spans, by a Kronecker product, first), so any
amplitude, probability or Pauli expectation of a 30- or 32-qubit register is a product of
small per-cluster quantities.  Nothing here is a restatement of a GPU kernel: the clusters
are advanced by the oracle, and the composition is the tensor-product identity
<i|psi> = prod_c <i restricted to c|phi_c>.

Test infrastructure only (imported by tests/).
"""
from __future__ import annotations

from typing import Dict, List, Sequence

import numpy as np

import oracle


class Factored:
    def __init__(self, n: int):
        self.n = n
        # cluster id -> (positions in local-bit order, complex128 vector over them)
        self.pos: Dict[int, List[int]] = {}
        self.vec: Dict[int, np.ndarray] = {}
        self.owner = list(range(n))
        for p in range(n):
            self.pos[p] = [p]
            self.vec[p] = np.array([1.0, 0.0], dtype=np.complex128)

    # ---- structure ----
    def _merge(self, positions: Sequence[int]) -> int:
        ids = []
        for p in positions:
            c = self.owner[p]
            if c not in ids:
                ids.append(c)
        head = ids[0]
        for c in ids[1:]:
            # new local bits: head's bits first, then c's (kron(b, a)[ib * Na + ia] = b[ib] a[ia])
            self.vec[head] = np.kron(self.vec[c], self.vec[head])
            self.pos[head] = self.pos[head] + self.pos[c]
            for p in self.pos[c]:
                self.owner[p] = head
            del self.pos[c], self.vec[c]
        if len(self.pos[head]) > 16:
            raise ValueError("cluster grew beyond 16 positions; the test circuit is too entangling")
        return head

    def _local(self, c: int, p: int) -> int:
        return self.pos[c].index(p)

    def cluster_sizes(self) -> List[int]:
        return sorted(len(v) for v in self.pos.values())

    # ---- gates (data positions, the C ABI's conventions) ----
    def apply_1q(self, p, m):
        c = self.owner[p]
        oracle.apply_1q(self.vec[c], self._local(c, p), m)

    def apply_c1q(self, cpos, t, m):
        c = self._merge([cpos, t])
        oracle.apply_c1q(self.vec[c], self._local(c, cpos), self._local(c, t), m)

    def apply_mc1q(self, controls, t, m):
        c = self._merge(list(controls) + [t])
        v = self.vec[c]
        k = len(self.pos[c])
        lt = self._local(c, t)
        cm = 0
        for q in controls:
            cm |= 1 << self._local(c, q)
        m = np.asarray(m, dtype=np.complex128)
        for i in range(1 << k):
            if (i & cm) == cm and not (i >> lt) & 1:
                j = i | (1 << lt)
                a, b = v[i], v[j]
                v[i] = m[0, 0] * a + m[0, 1] * b
                v[j] = m[1, 0] * a + m[1, 1] * b

    def apply_swaplike(self, p1, p2, m):
        # iqsh_apply_swaplike orders the pair first (as ApplySwap_helper does,
        # src/qureg_applyswap.cpp:137-144): m acts on (|lo = 1, hi = 0>, |lo = 0, hi = 1>).
        # The oracle names the pair by argument order, so it gets (lo, hi) whatever the
        # cluster's local bit order.
        lo, hi = min(p1, p2), max(p1, p2)
        c = self._merge([lo, hi])
        oracle.apply_swaplike(self.vec[c], self._local(c, lo), self._local(c, hi), m)

    def apply_diag2(self, p1, p2, d):
        c = self._merge([p1, p2])
        oracle.apply_diag(self.vec[c], self._local(c, p1), self._local(c, p2), d)

    def apply_2q(self, ph, pl, m):
        c = self._merge([ph, pl])
        oracle.apply_2q(self.vec[c], self._local(c, ph), self._local(c, pl), m)

    def apply_kq(self, positions, U):
        """Dense block, row index bit j = bit positions[j] (iqsh_apply_kq)."""
        c = self._merge(positions)
        v = self.vec[c]
        k = len(self.pos[c])
        loc = [self._local(c, p) for p in positions]
        U = np.asarray(U, dtype=np.complex128)
        kk = len(positions)
        # view v as a tensor with one axis per local bit (axis order: highest local bit first)
        t = v.reshape([2] * k)
        axes = [k - 1 - l for l in loc]          # tensor axis of local bit loc[j]
        # move the block's axes to the end, highest block bit first
        order = [a for a in range(k) if a not in axes] + [axes[j] for j in reversed(range(kk))]
        tt = np.transpose(t, order).reshape(-1, 1 << kk)
        tt = tt @ U.T
        inv = np.argsort(order)
        self.vec[c] = np.transpose(tt.reshape([2] * k), inv).reshape(-1).copy()

    def permute(self, src_pos):
        """New position b holds old position src_pos[b] (iqsh_permute)."""
        inv = [0] * self.n
        for b, s in enumerate(src_pos):
            inv[s] = b
        owner = [0] * self.n
        for c in self.pos:
            self.pos[c] = [inv[p] for p in self.pos[c]]
            for p in self.pos[c]:
                owner[p] = c
        self.owner = owner

    # ---- read-outs ----
    def amplitudes(self, idx: np.ndarray) -> np.ndarray:
        idx = np.asarray(idx, dtype=np.uint64)
        out = np.ones(idx.shape, dtype=np.complex128)
        for c, ps in self.pos.items():
            loc = np.zeros(idx.shape, dtype=np.int64)
            for j, p in enumerate(ps):
                loc |= (((idx >> np.uint64(p)) & np.uint64(1)).astype(np.int64)) << j
            out *= self.vec[c][loc]
        return out

    def norm2(self) -> float:
        r = 1.0
        for v in self.vec.values():
            r *= float(np.vdot(v, v).real)
        return r

    def prob1(self, p) -> float:
        c = self.owner[p]
        v = self.vec[c]
        l = self._local(c, p)
        sel = (np.arange(v.size) >> l) & 1 == 1
        r = float((np.abs(v[sel]) ** 2).sum())
        for cc, vv in self.vec.items():
            if cc != c:
                r *= float(np.vdot(vv, vv).real)
        return r

    def expval_pauli(self, xmask: int, zmask: int) -> float:
        """<psi| prod_{x only} X prod_{x and z} Y prod_{z only} Z |psi> (iqsh_expval_pauli)."""
        r = 1.0 + 0j
        for c, ps in self.pos.items():
            v = self.vec[c]
            k = len(ps)
            lx = lz = 0
            ny = 0
            for j, p in enumerate(ps):
                if (xmask >> p) & 1:
                    lx |= 1 << j
                if (zmask >> p) & 1:
                    lz |= 1 << j
                if (xmask >> p) & 1 and (zmask >> p) & 1:
                    ny += 1
            i = np.arange(1 << k)
            # P|i> = i^ny (-1)^popcount(i & lz) |i ^ lx>  (Y = i X Z)
            sign = 1 - 2 * (np.array([bin(int(a)).count("1") for a in (i & lz)]) & 1)
            pv = np.zeros_like(v)
            pv[i ^ lx] = (1j ** ny) * sign * v
            r *= np.vdot(v, pv)
        return float(r.real)
