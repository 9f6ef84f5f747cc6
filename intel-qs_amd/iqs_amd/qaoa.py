"""QAOA MaxCut on the GPU engine: forward pass, adjoint gradient, parameter-shift and
finite-difference references, and a torch.autograd.Function.

Circuit (the reference's form, src/qaoa_features.cpp:248-264 and its unit test's mixer
qaoa_features_test.hpp:75-76): start from |+>^n; layer l applies
    U_C(gamma_l) = exp(-i gamma_l C),  C = sum_e w_e (1 - Z_u Z_v) / 2   (the cut value)
    U_B(beta_l)  = prod_q RX_q(2 beta_l) = exp(-i beta_l sum_q X_q)
and the cost is <C> in the final state.

Adjoint method (no counterpart in the reference): after the forward pass psi = psi_p and
lambda = C psi.  Walking the gates backwards, each parameterized gate U = exp(-i theta G)
contributes dC/dtheta = 2 Im <lambda| G |psi> (psi is the state just after U), then U is
undone on both vectors.  The mixer's qubit rotations commute with one another and with
every X_q, so each <lambda|X_q|psi> can be taken in the fused pass that undoes its qubit's
rotation (six, then four qubits per pass: iqsh_adjoint_1q_all).  Memory: the state,
lambda, and nothing else (no checkpoints).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Sequence, Tuple

import numpy as np

from . import gates as G
from .circuits import SplitMix64, random_regular_graph
from .engine import StateVector

X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
# Registers from this size on apply the mixer with the fused layer kernels
# (iqsh_apply_1q_all / iqsh_adjoint_1q_all); smaller ones qubit by qubit.
LAYER_MIN_QUBITS = 10


@dataclass
class MaxCutQaoa:
    num_qubits: int
    edges: List[Tuple[int, int, float]]
    dtype: str = "c128"
    renorm_every: int = 0  # complex64: renormalize (FP64 norm) every k layers; 0 = never
    _psi: StateVector = field(default=None, init=False, repr=False)
    _lam: StateVector = field(default=None, init=False, repr=False)

    @staticmethod
    def random_3regular(n: int, seed: int, dtype: str = "c128") -> "MaxCutQaoa":
        return MaxCutQaoa(n, random_regular_graph(n, 3, seed), dtype)

    def _state(self, which: str) -> StateVector:
        sv = getattr(self, which)
        if sv is None:
            sv = StateVector(self.num_qubits, self.dtype)
            setattr(self, which, sv)
        return sv

    def _prepare(self, psi: StateVector) -> None:
        psi.init_uniform(1.0 / math.sqrt(float(1 << self.num_qubits)))

    def _mixer(self, psi: StateVector, beta: float) -> None:
        m = G.rx(2.0 * beta)
        if self.num_qubits >= LAYER_MIN_QUBITS:
            psi.apply_1q_all(m)  # RX on every qubit in fused passes
            return
        for q in range(self.num_qubits):
            psi.apply_1q(q, m)

    def _renorm(self, psi: StateVector) -> None:
        psi.scale(1.0 / math.sqrt(psi.norm2()))

    def forward_state(self, gammas: Sequence[float], betas: Sequence[float], shift=None) -> StateVector:
        """Run the circuit; `shift` = (kind, layer, index, delta) adds one extra rotation
        (an edge phase or a qubit RX) used by the parameter-shift rule."""
        psi = self._state("_psi")
        self._prepare(psi)
        fused = self.num_qubits >= LAYER_MIN_QUBITS
        for l, (g, b) in enumerate(zip(gammas, betas)):
            if shift and shift[0] == "edge" and shift[1] == l:
                # diagonal phases commute: the extra edge phase may precede the layer
                psi.qaoa_phase([self.edges[shift[2]]], shift[3])
            if fused:
                psi.qaoa_layer(self.edges, g, G.rx(2.0 * b))  # phase fused into the mixer
            else:
                psi.qaoa_phase(self.edges, g)
                self._mixer(psi, b)
            if shift and shift[0] == "qubit" and shift[1] == l:
                psi.apply_1q(shift[2], G.rx(2.0 * shift[3]))
            if self.renorm_every and (l + 1) % self.renorm_every == 0:
                self._renorm(psi)
        return psi

    def cost(self, gammas, betas, shift=None) -> float:
        return self.forward_state(gammas, betas, shift).qaoa_expect(self.edges)

    def adjoint(self, gammas: Sequence[float], betas: Sequence[float]) -> Tuple[float, np.ndarray, np.ndarray]:
        """(cost, dC/dgamma, dC/dbeta) by the adjoint method."""
        self.forward_state(gammas, betas)
        return self.adjoint_from_final(gammas, betas)

    def adjoint_from_final(self, gammas, betas) -> Tuple[float, np.ndarray, np.ndarray]:
        """Adjoint sweep starting from the final state left by forward_state()."""
        p = len(gammas)
        psi = self._state("_psi")
        lam = self._state("_lam")
        cost = psi.qaoa_cost_apply(lam, self.edges)
        dg, db = np.zeros(p), np.zeros(p)
        for l in reversed(range(p)):
            uinv = G.rx(-2.0 * betas[l])
            if self.num_qubits >= LAYER_MIN_QUBITS:
                dg[l], db[l] = psi.qaoa_layer_adjoint(lam, self.edges, gammas[l], uinv, X)
                continue
            s = 0.0
            for q in reversed(range(self.num_qubits)):
                s += psi.adjoint_1q(lam, q, uinv, X)
            db[l] = s
            dg[l] = psi.qaoa_phase_adjoint(lam, self.edges, gammas[l])
        return cost, dg, db

    def parameter_shift(self, gammas, betas, which: Sequence[Tuple[str, int]] = None):
        """Exact gradient by the two-term shift rule.  Each edge term w (1 - ZZ)/2 has
        eigenvalues {0, w}: d/dgamma = (w/2) [f(+pi/(2w)) - f(-pi/(2w))] per edge (pi/2 shift,
        factor 1/2 for unit weights); each X_q in the mixer exp(-i beta X_q) has
        eigenvalues +-1: d/dbeta = f(+pi/4) - f(-pi/4) per qubit.  `which` selects
        parameters as ("gamma", l) / ("beta", l); default all.  Returns a dict and the
        number of circuit evaluations."""
        p = len(gammas)
        which = which or [("gamma", l) for l in range(p)] + [("beta", l) for l in range(p)]
        out, nevals = {}, 0
        for kind, l in which:
            g = 0.0
            if kind == "gamma":
                for e, (_, _, w) in enumerate(self.edges):
                    d = math.pi / (2.0 * w)
                    g += 0.5 * w * (self.cost(gammas, betas, ("edge", l, e, d)) -
                                    self.cost(gammas, betas, ("edge", l, e, -d)))
                    nevals += 2
            else:
                for q in range(self.num_qubits):
                    g += (self.cost(gammas, betas, ("qubit", l, q, math.pi / 4)) -
                          self.cost(gammas, betas, ("qubit", l, q, -math.pi / 4)))
                    nevals += 2
            out[(kind, l)] = g
        return out, nevals

    def finite_difference(self, gammas, betas, kind: str, l: int, h: float = 1e-4) -> float:
        gp, gm = list(gammas), list(gammas)
        bp, bm = list(betas), list(betas)
        if kind == "gamma":
            gp[l] += h
            gm[l] -= h
        else:
            bp[l] += h
            bm[l] -= h
        return (self.cost(gp, bp) - self.cost(gm, bm)) / (2.0 * h)


def default_problem(n: int = 28, p: int = 8, dtype: str = "c128"):
    """3-regular MaxCut graph (seed 20240604) and (gamma, beta) ~ U[0, pi] (seed 20240605)."""
    prob = MaxCutQaoa.random_3regular(n, 20240604, dtype)
    rng = SplitMix64(20240605)
    gammas = [rng.uniform(0.0, math.pi) for _ in range(p)]
    betas = [rng.uniform(0.0, math.pi) for _ in range(p)]
    return prob, gammas, betas


def torch_function():
    """torch.autograd.Function computing <C>(gammas, betas) with the adjoint backward."""
    import torch

    class QaoaExpectation(torch.autograd.Function):
        """forward: one circuit pass, <C> returned; the final state stays on the GPU.
        backward: lambda = C psi and the adjoint sweep from that state (the forward is
        re-run only if the problem object ran another circuit in between)."""

        @staticmethod
        def forward(ctx, gammas, betas, problem):
            g = gammas.detach().cpu().double().numpy()
            b = betas.detach().cpu().double().numpy()
            cost = problem.forward_state(g, b).qaoa_expect(problem.edges)
            ctx.token = object()
            problem._token = ctx.token
            ctx.problem, ctx.params = problem, (g, b)
            ctx.devices = (gammas.device, betas.device)
            return torch.tensor(cost, dtype=torch.float64, device=gammas.device)

        @staticmethod
        def backward(ctx, grad_out):
            prob, (g, b) = ctx.problem, ctx.params
            if getattr(prob, "_token", None) is not ctx.token:
                prob.forward_state(g, b)
            _, dg, db = prob.adjoint_from_final(g, b)
            prob._token = None
            go = grad_out.detach().double().cpu()
            return ((go * torch.from_numpy(dg)).to(ctx.devices[0]), (go * torch.from_numpy(db)).to(ctx.devices[1]),
                    None)

    return QaoaExpectation
