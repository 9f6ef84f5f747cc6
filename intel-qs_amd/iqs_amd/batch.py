"""Batches of hardware-efficient-ansatz (HEA) circuits on the GPU engine.

Instance b has n_b qubits and d_b layers and starts from |0...0>.  Layer l applies, on
every qubit q, ApplyRotationY(q, theta[l][q]) then ApplyRotationZ(q, phi[l][q])
(src/qureg_apply1qubitgate.cpp:297-325), then ApplyCPauliZ(i, i+1) for i = 0..n_b-2
(src/qureg_applyctrl1qubitgate.cpp:556).  The cost is sum_q <Z_q>
(ExpectationValueZ, src/qureg_expectval.cpp:70).  The whole batch -- every forward pass
and the adjoint gradient of every angle -- runs in one persistent kernel launch replayed
from a HIP graph (libiqs_hip.so, iqsh_hea_*).

Parameter layout (angles and gradients): instance after instance, then layer, qubit and
the pair (theta, phi): 2 * n_b * d_b doubles per instance.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._lib import HEA_GRAD, HEA_KEEP_STATE, HEA_NO_GRAPH, HeaTrial, HeaTune, call, lib
from .circuits import Circuit, SplitMix64

DTYPES = {"c64": _lib.C64, "c128": _lib.C128}


def random_batch(count: int = 512, seed: int = 20240606, qubits=(10, 22), depth=(10, 50)):
    """(nqubits, depths, angles) of a seeded batch: n_b ~ U{qubits}, d_b ~ U{depth} drawn
    instance by instance, then every angle ~ U[0, 2pi) in parameter order (SplitMix64)."""
    rng = SplitMix64(seed)
    nq, dp = [], []
    for _ in range(count):
        nq.append(qubits[0] + rng.randint(qubits[1] - qubits[0] + 1))
        dp.append(depth[0] + rng.randint(depth[1] - depth[0] + 1))
    npar = 2 * sum(n * d for n, d in zip(nq, dp))
    ang = np.array([rng.uniform(0.0, 2.0 * math.pi) for _ in range(npar)], dtype=np.float64)
    return nq, dp, ang


def param_offsets(nqubits: Sequence[int], depths: Sequence[int]) -> np.ndarray:
    """Offset (in doubles) of each instance's parameters, plus the total at the end."""
    off = [0]
    for n, d in zip(nqubits, depths):
        off.append(off[-1] + 2 * n * d)
    return np.array(off, dtype=np.int64)


def instance_circuit(n: int, d: int, params: np.ndarray, precision: str = "dp", readout: bool = True) -> Circuit:
    """Instance as a gate list of QubitRegister method names (for the CPU oracle)."""
    p = np.asarray(params, dtype=np.float64).reshape(d, n, 2)
    c = Circuit(n, precision=precision)
    for l in range(d):
        for q in range(n):
            c.add("ApplyRotationY", q, float(p[l, q, 0]))
            c.add("ApplyRotationZ", q, float(p[l, q, 1]))
        for i in range(n - 1):
            c.add("ApplyCPauliZ", i, i + 1)
    if readout:
        c.add("ZAll")
    return c


class HeaBatch:
    """A batch of HEA instances resident on the GPU (states, matrices, work buffers)."""

    def __init__(self, nqubits: Sequence[int], depths: Sequence[int], dtype: str = "c64",
                 tune: Optional[Tuple[int, ...]] = None):
        self.nqubits = [int(x) for x in nqubits]
        self.depths = [int(x) for x in depths]
        self.count = len(self.nqubits)
        self.dtype = dtype
        self.offsets = param_offsets(self.nqubits, self.depths)
        n_arr = (C.c_int * self.count)(*self.nqubits)
        d_arr = (C.c_int * self.count)(*self.depths)
        t = HeaTune(*tune) if tune is not None else None
        h = C.c_void_p()
        call("iqsh_hea_create", DTYPES[dtype], self.count, n_arr, d_arr, C.byref(t) if t else None, C.byref(h))
        self.h = h
        npar = C.c_uint64()
        call("iqsh_hea_num_params", self.h, C.byref(npar))
        self.num_params = int(npar.value)
        assert self.num_params == int(self.offsets[-1])

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().iqsh_hea_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- tuning ----
    @staticmethod
    def variants(dtype: str = "c64") -> List[dict]:
        cnt = C.c_int()
        call("iqsh_hea_variants", DTYPES[dtype], C.byref(cnt))
        out = []
        for v in range(cnt.value):
            m, r, t, ft = C.c_int(), C.c_int(), C.c_int(), C.c_int()
            call("iqsh_hea_variant_info", DTYPES[dtype], v, C.byref(m), C.byref(r), C.byref(t), C.byref(ft))
            out.append(dict(variant=v, tile_qubits=m.value, amps_per_thread=r.value, threads=t.value,
                            for_tests=bool(ft.value)))
        return out

    @property
    def tune(self) -> Tuple[int, int, int, int]:
        """(variant, workers_per_cu, order, group_mib), see iqsh_hea_tune."""
        t = HeaTune()
        call("iqsh_hea_get_tune", self.h, C.byref(t))
        return (t.variant, t.workers_per_cu, t.order, t.group_mib)

    @tune.setter
    def tune(self, value: Tuple[int, ...]) -> None:
        t = HeaTune(*value)
        call("iqsh_hea_set_tune", self.h, C.byref(t))

    def autotune(self, grad: bool = True, reps: int = 2):
        """Time every tuning on this batch and keep the fastest.  Returns (tune, ms, trials)."""
        trials = (HeaTrial * 64)()
        nt, best = C.c_int(), C.c_double()
        call("iqsh_hea_autotune", self.h, HEA_GRAD if grad else 0, reps, trials, 64, C.byref(nt), C.byref(best))
        tr = [((x.tune.variant, x.tune.workers_per_cu, x.tune.order, x.tune.group_mib), x.ms)
              for x in trials[:min(nt.value, 64)]]
        return self.tune, best.value, tr

    # ---- data ----
    def set_angles(self, angles: np.ndarray) -> None:
        a = np.ascontiguousarray(angles, dtype=np.float64)
        if a.size != self.num_params:
            raise ValueError(f"expected {self.num_params} angles, got {a.size}")
        call("iqsh_hea_set_angles", self.h, a.ctypes.data)

    def set_angles_device(self, ptr: int) -> None:
        call("iqsh_hea_set_angles_device", self.h, C.c_void_p(ptr))

    def run(self, grad: bool = True, keep_state: bool = False, graph: bool = True) -> None:
        flags = (HEA_GRAD if grad else 0) | (HEA_KEEP_STATE if keep_state else 0) | (0 if graph else HEA_NO_GRAPH)
        call("iqsh_hea_run", self.h, flags)

    def results(self, grad: bool = True):
        cost = np.zeros(self.count, dtype=np.float64)
        g = np.zeros(self.num_params, dtype=np.float64) if grad else None
        call("iqsh_hea_results", self.h, cost.ctypes.data, g.ctypes.data if grad else None)
        return (cost, g) if grad else cost

    def results_device(self, cost_ptr: int, grad_ptr: Optional[int]) -> None:
        """Copy costs and gradients into device memory; raises (copying nothing) if the run
        aborted or left a task queue undrained."""
        call("iqsh_hea_results_device", self.h, C.c_void_p(cost_ptr), C.c_void_p(grad_ptr) if grad_ptr else None)

    def set_wait_limit(self, max_polls: int = 1 << 20, unmet: bool = False) -> None:
        """Bounded dependency waits: polls before a tile wait gives up and the run reports an
        error; unmet=True makes every wait unsatisfiable (fault injection for tests)."""
        call("iqsh_hea_set_wait_limit", self.h, int(max_polls), int(bool(unmet)))

    def state(self, b: int) -> np.ndarray:
        dt = np.complex64 if self.dtype == "c64" else np.complex128
        st = np.zeros(1 << self.nqubits[b], dtype=dt)
        call("iqsh_hea_state", self.h, int(b), st.ctypes.data)
        return st

    def traffic(self, grad: bool = True, keep_state: bool = False) -> float:
        b = C.c_double()
        call("iqsh_hea_traffic", self.h, (HEA_GRAD if grad else 0) | (HEA_KEEP_STATE if keep_state else 0), C.byref(b))
        return b.value

    # ---- convenience ----
    def evaluate(self, angles: np.ndarray, grad: bool = True):
        self.set_angles(angles)
        self.run(grad=grad)
        return self.results(grad=grad)

    def parameter_shift(self, angles: np.ndarray, slots: Sequence[int]) -> np.ndarray:
        """Exact d cost_b / d angle for one parameter slot per instance (slots[b] = index
        within instance b's parameters) by the two-term shift rule: RY and RZ are
        exp(-i a P/2) with P^2 = 1, so dC/da = [C(a + pi/2) - C(a - pi/2)] / 2.  All
        instances are shifted together: two forward-only batch runs."""
        idx = self.offsets[:-1] + np.asarray(slots, dtype=np.int64)
        a = np.array(angles, dtype=np.float64)
        a[idx] += 0.5 * math.pi
        cp = self.evaluate(a, grad=False)
        a[idx] -= math.pi
        cm = self.evaluate(a, grad=False)
        return 0.5 * (cp - cm)


def torch_function():
    """torch.autograd.Function: costs (count,) = f(angles (num_params,)) for a HeaBatch.
    The forward launch already computes the adjoint gradient; backward scales it by the
    incoming per-instance gradient."""
    import torch

    class HeaCost(torch.autograd.Function):
        @staticmethod
        def forward(ctx, angles, batch):
            dev = angles.device
            a = angles.detach().to(torch.float64).contiguous()
            if a.is_cuda:
                torch.cuda.current_stream(dev).synchronize()
                batch.set_angles_device(a.data_ptr())
            else:
                batch.set_angles(a.numpy())
            batch.run(grad=True)
            cost = torch.empty(batch.count, dtype=torch.float64, device=dev)
            grad = torch.empty(batch.num_params, dtype=torch.float64, device=dev)
            if dev.type == "cuda":
                torch.cuda.current_stream(dev).synchronize()
                batch.results_device(cost.data_ptr(), grad.data_ptr())
                _lib.call("iqsh_sync")
            else:
                c, g = batch.results(grad=True)
                cost.copy_(torch.from_numpy(c))
                grad.copy_(torch.from_numpy(g))
            ctx.save_for_backward(grad)
            ctx.counts = np.diff(batch.offsets)
            return cost

        @staticmethod
        def backward(ctx, grad_out):
            (grad,) = ctx.saved_tensors
            reps = torch.as_tensor(ctx.counts, device=grad.device)
            scale = torch.repeat_interleave(grad_out.to(torch.float64), reps)
            return grad * scale, None

    return HeaCost
