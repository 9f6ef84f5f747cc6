#!/usr/bin/env python3
"""Gate-throughput benchmark of the MI355X state-vector engine.

Workload (one GPU): a 30-qubit ComplexDP register (16 GiB in HBM) running the random brick
circuit of depth 40 (seed 20240602): per layer, H on every qubit, R{X,Y,Z}(theta) on every
qubit, CNOT(i, i+1) for i = layer mod 2, +2, ...  One step = one layer.  `value` is gates/s
over the timed layers with the state resident in HBM.

Also reported:
  roofline      the dominant kernel (dense 1q update, used by H/RX/RY) timed per launch with
                HIP events on the library stream inside the timed region; algorithmic bytes
                per launch = 2 * 2^n * 16 B (read + write of the whole state)
  sweep         GB/s per gate kind at every target position (same byte model as the
                reference's Timer bookkeeping, src/highperfkernels.cpp:230-231, 370, 478-479)
  cpu_baseline  the compiled reference (oracle/_ref/ref_driver, OpenMP) on a bounded sample
                of the same circuit on this host's cores
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "intel-qs_amd"))

METRIC = "30q gate/s & %HBM peak; QAOA-28q grad s: adjoint vs param-shift @1e-6"
HBM_PEAK_GBS = 8000.0


_T0 = time.perf_counter()


def log(msg: str) -> None:
    """Progress line on stderr (the JSON result is the only stdout line)."""
    print(f"bench.py [{time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def apply_op(sv, op, G):
    name = op[0]
    if name in G.NAMED_1Q:
        sv.apply_1q(op[1], G.NAMED_1Q[name]())
    elif name in G.ROT_1Q:
        sv.apply_1q(op[1], G.ROT_1Q[name](op[2]))
    elif name in G.NAMED_C:
        sv.apply_c1q(op[1], op[2], G.NAMED_C[name]())
    else:
        raise ValueError(name)


def apply_inverse_op(sv, op, G):
    """The inverse of apply_op's gate: H and CNOT are involutions, R(theta)^-1 = R(-theta)."""
    name = op[0]
    if name in G.ROT_1Q:
        sv.apply_1q(op[1], G.ROT_1Q[name](-op[2]))
    else:
        apply_op(sv, op, G)


def host_info() -> dict:
    """CPU model, logical CPUs and physical cores of this host (lscpu)."""
    info = {"nproc": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=20).stdout
        cores = {tuple(l.split(",")) for l in out.splitlines() if l and not l.startswith("#")}
        info["physical_cores"] = len(cores)
        for l in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout.splitlines():
            if l.startswith("Model name:"):
                info["model"] = l.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def kernel_kind(op):
    """Which kernel a brick-circuit gate launches (iqs_capi.hip dispatch)."""
    name = op[0]
    if name in ("ApplyHadamard", "ApplyRotationX", "ApplyRotationY"):
        return "dense1q_pair" if op[1] >= 6 else "dense1q_lane"
    if name == "ApplyRotationZ":
        return "diag1q"
    return "cnot"


def traffic_from_profile(n: int):
    """HBM bytes per launch of the dominant kernel, from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by the copy-kernel calibration)."""
    path = os.path.join(ROOT, "profiles", "k_pair_dense1q_30q_pmc_traffic.md")
    if n != 30 or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def profile_value(name: str, key: str):
    """A `key: value` line of a committed profiles/ summary, or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":")[1])
    return None


def calibrate(n: int) -> int:
    """Run one device-to-device copy of a known size (for PMC counter calibration)."""
    from iqs_amd.engine import StateVector, sync
    a, b = StateVector(n, "c128"), StateVector(n, "c128")
    a.init_uniform(0.5)
    sync()
    b.copy_from(a)
    sync()
    print(f"calibration copy: {a.size * 16} bytes read and {a.size * 16} bytes written")
    return 0


def _time_ms(fn, reps, stream, torch):
    """Average milliseconds of fn() over reps calls (one warm-up), HIP events on `stream`."""
    fn()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(reps):
        fn()
    b.record(stream)
    b.synchronize()
    return a.elapsed_time(b) / reps


def _rates(rows, nbytes):
    ts = [t for _, t in rows]
    return {"bytes_per_call": nbytes, "avg_ms": sum(ts) / len(ts),
            "min_GBs": nbytes / max(ts) / 1e6, "avg_GBs": nbytes / (sum(ts) / len(ts)) / 1e6,
            "max_GBs": nbytes / min(ts) / 1e6, "per_case_GBs": [round(nbytes / t / 1e6, 1) for _, t in rows]}


def kind_sweep(sv, n: int, esz: int, reps: int, stream, torch):
    """GB/s per gate kind at every target position (the reference's byte model, Timer
    bookkeeping of src/highperfkernels.cpp:230-231, 370, 478-479: a dense update reads and
    writes the state, 2S; a controlled gate or a swap touches half of it, S)."""
    import numpy as np
    from iqs_amd import gates as G
    from iqs_amd.circuits import SplitMix64, haar_unitary, random_su2

    S = (1 << n) * esz
    srng = SplitMix64(7)
    kinds = {
        "1q_dense": ([(p,) for p in range(n)], lambda p: sv.apply_1q(p[0], random_su2(srng)), 2 * S),
        "controlled_1q": ([((t + 1) % n, t) for t in range(n)] + [(0, n - 1), (n - 1, 0)],
                          lambda p: sv.apply_c1q(p[0], p[1], random_su2(srng)), S),
        "diag_2q": ([(p, (p + 1) % n) for p in range(n)],
                    lambda p: sv.apply_diag2(p[0], p[1], np.exp(1j * np.array([srng.uniform(0, 6) for _ in range(4)]))),
                    2 * S),
        "swap": ([(p, n - 1 - p) for p in range(n // 2)], lambda p: sv.apply_swaplike(p[0], p[1], G.SWAP), S),
        "dense_2q": ([(p, p + 1) for p in range(n - 1)] + [(0, n - 1)],
                     lambda p: sv.apply_2q(p[0], p[1], haar_unitary(4, srng)), 2 * S),
    }
    out = {}
    for kname, (targets, fn, nbytes) in kinds.items():
        rows = [(t, _time_ms(lambda: fn(t), reps, stream, torch)) for t in targets]
        out[kname] = dict(_rates(rows, nbytes), cases=[list(t) for t in targets])
        log(f"sweep {kname} ({n} qubits, {esz} B): {out[kname]['avg_GBs']:.0f} GB/s average")
    return out


def pool_sweep(reps: int, stream, torch):
    """Pool of states (the reference's pool, src/mpi_env.cpp:239-404, as one register whose
    top positions index the states): 8 states of 27 qubits complex128 and 8 of 28 complex64
    (16 GiB each), per-state 1q (noise gates, 1q channel branches) at low, middle and top
    positions and per-state 2q (2q channel branches), against 2S; per-state norm, P(qubit)
    and Z-string sums against S; per-state overlap with a reference state against 2S (the
    pool and the reference once per state).  Every call includes its host work: the matrix
    upload of the per-state gates, the per-state sums of the reductions."""
    import ctypes as C

    import numpy as np
    from iqs_amd._lib import call
    from iqs_amd.circuits import SplitMix64, haar_unitary
    from iqs_amd.engine import StateVector

    out = {}
    for dt, sq in (("c128", 27), ("c64", 28)):
        ns = 8
        n = sq + 3
        S = (1 << n) * (16 if dt == "c128" else 8)
        sv = StateVector(n, dt)
        call("iqsh_pool_init", sv.handle, sq, 1, 0, ns)
        rng = SplitMix64(20240605)

        def mats(k):
            return np.concatenate([np.stack([u.real, u.imag], -1).reshape(-1) for u in
                                   (haar_unitary(1 << k, rng) for _ in range(ns))]).astype(np.float64)

        m2, m4 = mats(1), mats(2)
        p2, p4 = (m.ctypes.data_as(C.POINTER(C.c_double)) for m in (m2, m4))
        res = {}
        rows = [((p,), _time_ms(lambda: call("iqsh_pool_apply_1q_each", sv.handle, sq, p, p2, ns), reps, stream, torch))
                for p in (0, 2, 5, 13, sq - 1)]
        res["per_state_1q"] = dict(_rates(rows, 2 * S), cases=[list(t) for t, _ in rows])
        rows = [((a, b), _time_ms(lambda: call("iqsh_pool_apply_2q_each", sv.handle, sq, a, b, p4, ns), reps, stream,
                                  torch)) for a, b in ((1, 0), (5, 20), (sq - 1, sq - 2))]
        res["per_state_2q"] = dict(_rates(rows, 2 * S), cases=[list(t) for t, _ in rows])
        red = (C.c_double * (2 * ns))()
        rows = [(nm, _time_ms(lambda: call("iqsh_pool_reduce", sv.handle, sq, kind, arg, red, ns), reps, stream, torch))
                for nm, kind, arg in (("norm", 0, 0), ("prob1_3", 1, 3), ("zstring", 2, (1 << 5) | 1))]
        res["per_state_reductions"] = dict(_rates(rows, S), cases=[t for t, _ in rows])
        ref = StateVector(sq, dt)
        ref.init_basis(0)
        ms = _time_ms(lambda: call("iqsh_pool_overlap", sv.handle, sq, ref.handle, red, ns), reps, stream, torch)
        res["per_state_overlap_GBs"] = 2 * S / ms / 1e6
        ref.close()
        sv.close()
        res.update({"states": ns, "state_qubits": sq})
        out[dt] = res
        log(f"sweep pool ({ns} x {sq} qubits, {dt}): per-state 1q {res['per_state_1q']['avg_GBs']:.0f} GB/s average")
    return out


def permute_sweep(sv, n: int, esz: int, reps: int, stream, torch):
    """PermuteQubits (src/qureg_permute.cpp:57-105) as one bit-exact index permutation:
    reads and writes the state once (2S).  Each case is applied and then undone, so the
    register ends where it started."""
    import numpy as np

    S = (1 << n) * esz
    cases = {
        "reversal": list(reversed(range(n))),
        "rotate_up_1": [(b + 1) % n for b in range(n)],
        "rotate_down_1": [(b - 1) % n for b in range(n)],
        "swap_halves": [(b + n // 2) % n for b in range(n)],
        "swap_0_top": [n - 1] + list(range(1, n - 1)) + [0],
        "random": [int(x) for x in np.random.default_rng(n).permutation(n)],
    }
    rows = []
    for name, perm in cases.items():
        inv = [0] * n
        for b, src in enumerate(perm):
            inv[src] = b
        ms = _time_ms(lambda: (sv.permute(perm), sv.permute(inv)), reps, stream, torch) / 2
        rows.append((name, ms))
    out = _rates(rows, 2 * S)
    out["cases"] = {k: v for k, v in cases.items()}
    log(f"sweep PermuteQubits ({n} qubits, {esz} B): {out['avg_GBs']:.0f} GB/s average")
    return out


def reduction_sweep(sv, n: int, esz: int, reps: int, stream, torch):
    """FP64 reductions: GetProbability at every position (reads the half with the bit set,
    S/2; src/qureg_measure.cpp:136-178), the norm (S) and one-pass Pauli expectations with
    X, Y and Z factors (S; replaces src/qureg_expectval.cpp:102-213)."""
    S = (1 << n) * esz
    prob = [(p, _time_ms(lambda: sv.prob1(p), reps, stream, torch)) for p in range(n)]
    H = n - 1
    paulis = {"Z0": (0, 1), "ZtopZ0": (0, 1 | (1 << H)), "X0": (1, 0), "Xtop": (1 << H, 0),
              "Y0Y1": (3, 3), "X0Ztop": (1, 1 << H), "XYZ_spread": ((1 << 2) | (1 << (H - 3)), (1 << (H - 3)) | (1 << 7))}
    pv = [(k, _time_ms(lambda: sv.expval_pauli(*m), reps, stream, torch)) for k, m in paulis.items()]
    out = {"prob1": _rates(prob, S // 2),
           "norm2": _rates([("norm", _time_ms(sv.norm2, reps, stream, torch))], S),
           "expval_pauli": dict(_rates(pv, S), cases=list(paulis))}
    log(f"sweep reductions ({n} qubits, {esz} B): prob1 {out['prob1']['avg_GBs']:.0f}, "
        f"expval {out['expval_pauli']['avg_GBs']:.0f} GB/s average")
    return out


def cpu_baseline(n: int, layer_ops, budget_s: float = 30.0):
    """Time the compiled reference on layer 0 of the same circuit (bounded by budget_s)."""
    drv = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    if not os.path.exists(drv):
        return None
    from iqs_amd.circuits import Circuit
    c = Circuit(n)
    for op in layer_ops:
        c.add(*op)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    with tempfile.TemporaryDirectory() as d:
        cf = os.path.join(d, "c.txt")
        c.write(cf)
        try:
            r = subprocess.run([drv, cf, "--time"], capture_output=True, text=True, timeout=budget_s * 6,
                               env=dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="spread",
                                        OMP_PLACES="cores"))
        except subprocess.TimeoutExpired:
            return None
    times = [float(l.split()[3]) for l in r.stderr.splitlines() if l.startswith("T ")]
    if not times:
        return None
    return {"value": len(times) / sum(times), "unit": "gates/s", "cores": threads, "kind": "reference",
            "sample": f"{n}-qubit ComplexDP brick layer 0 ({len(times)} gates) through the compiled reference "
                      f"(intel-qs src/, -O3 OpenMP, OMP_PROC_BIND=spread OMP_PLACES=cores), per-gate steady_clock",
            "host": host_info()}


def cpu_kind_sample(n: int):
    """The per-gate-kind sweep's byte model on the compiled reference at n qubits: positions
    low, middle and high per kind, each gate applied twice (the first run is a warm-up),
    except the reference's serial kinds (ApplyDiag, Apply2QubitGate run on one thread, about
    1 GB/s): one middle position, applied once.  GB/s from the driver's per-gate timer."""
    import numpy as np
    from iqs_amd.circuits import Circuit, SplitMix64, haar_unitary, random_su2

    drv = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    if not os.path.exists(drv):
        return None
    r = SplitMix64(11)
    H, M = n - 1, n // 2
    S = (1 << n) * 16
    kinds = [
        ("1q_dense", 2 * S, 2, [("Apply1QubitGate", p, random_su2(r)) for p in (0, M, H)]),
        ("controlled_1q", S, 2, [("ApplyControlled1QubitGate", c, t, random_su2(r)) for c, t in ((1, 0), (M + 1, M), (0, H))]),
        ("swap", S, 2, [("ApplySwap", a, b) for a, b in ((0, H), (7, H - 7), (M - 1, M))]),
        ("diag_2q", 2 * S, 1, [("ApplyDiag", M, M + 1, np.exp(1j * np.array([r.uniform(0, 6) for _ in range(4)])))]),
        ("dense_2q", 2 * S, 1, [("Apply2QubitGate", M + 1, M, haar_unitary(4, r))]),
    ]
    c = Circuit(n)
    plan = []
    for name, nbytes, rep, ops in kinds:
        for op in ops:
            for k in range(rep):
                c.add(*op)
                plan.append((name, nbytes, k == rep - 1))
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    with tempfile.TemporaryDirectory() as d:
        cf = os.path.join(d, "c.txt")
        c.write(cf)
        try:
            res = subprocess.run([drv, cf, "--time"], capture_output=True, text=True, timeout=600,
                                 env=dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="spread",
                                          OMP_PLACES="cores"))
        except subprocess.TimeoutExpired:
            return None
    times = [float(l.split()[3]) for l in res.stderr.splitlines() if l.startswith("T ")]
    if len(times) != len(plan):
        return None
    out = {}
    for (name, nbytes, timed), t in zip(plan, times):
        if timed:
            out.setdefault(name, []).append(round(nbytes / t / 1e9, 2))
    return {"qubits": n, "threads": threads, "GBs_per_case": out,
            "cases": "positions low / middle / high per kind; ApplyDiag and Apply2QubitGate (serial in the "
                     "reference) at the middle position only"}


def brick20_against_reference(torch):
    """The reference's CPU-runnable case: a 20-qubit ComplexDP brick circuit of depth 20
    (seed 20240601) with <Z_i> read-outs, run through the drop-in on the GPU and through the
    compiled reference (oracle/_ref/ref_driver, OpenMP) on this host.  Reports both wall
    times and how far the results differ (the reference is the checker here)."""
    import numpy as np
    from iqs_amd import runner
    from iqs_amd.circuits import brick_circuit

    drv = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    if not os.path.exists(drv):
        return None
    c = brick_circuit(20, 20, 20240601)
    runner.run(c)  # warm-up (allocations, first launches)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reg, out = runner.run(c)
    gpu_state = np.array(reg, copy=True)
    gpu_s = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    with tempfile.TemporaryDirectory() as d:
        cf, sf, rf = (os.path.join(d, x) for x in ("c.txt", "s.bin", "r.txt"))
        c.write(cf)
        t0 = time.perf_counter()
        r = subprocess.run([drv, cf, "--state", sf, "--results", rf, "--time"], capture_output=True, text=True,
                           timeout=600, env=dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="spread",
                                                 OMP_PLACES="cores"))
        cpu_wall = time.perf_counter() - t0
        if r.returncode != 0:
            return None
        ref_state = np.fromfile(sf, dtype=np.complex128)
        ref_out = [l for l in open(rf).read().splitlines() if l]
    gate_s = sum(float(l.split()[3]) for l in r.stderr.splitlines() if l.startswith("T "))
    zg = [float(x) for l in out for x in l.split()[2:]]
    zr = [float(x) for l in ref_out for x in l.split()[2:]]
    return {"workload": "20-qubit ComplexDP brick circuit, depth 20, seed 20240601, <Z_i> read-outs",
            "gates": len(c.ops) - 1, "gpu_s": gpu_s, "cpu_gate_s": gate_s, "cpu_wall_s": cpu_wall, "cpu_threads": threads,
            "max_abs_amplitude_diff": float(np.max(np.abs(gpu_state - ref_state))),
            "max_abs_z_diff": float(max(abs(a - b) for a, b in zip(zg, zr))) if zg and len(zg) == len(zr) else None}


def qaoa_param_shift_on_reference(torch, n: int = 16, p: int = 8):
    """QAOA MaxCut gradient at 16 qubits, p = 8 (3-regular graph seed 20240604, angles seed
    20240605): parameter-shift through the compiled reference on this host (one driver run
    holding every shifted circuit: 2 evaluations per edge and per qubit and layer, the
    reference has no gradient code) against the GPU adjoint on the same problem."""
    import numpy as np
    from iqs_amd.circuits import Circuit
    from iqs_amd.qaoa import default_problem

    drv = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    if not os.path.exists(drv):
        return None
    prob, g, b = default_problem(n, p, "c128")
    prob.adjoint(g, b)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cost, dg, db = prob.adjoint(g, b)
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    c = Circuit(n, precision="dp", init="plus", graph=list(prob.edges))
    plan = []  # (kind, layer, coefficient) per pair of evaluations

    def evaluation(shift):
        c.add("Initialize", "++++", 0)
        for l, (gl, bl) in enumerate(zip(g, b)):
            c.add("QaoaPhase", float(gl))
            if shift[0] == "edge" and shift[1] == l:
                u, v, w = prob.edges[shift[2]]
                ph = -shift[3] * w
                c.add("ApplyDiag", u, v, 1.0, 0.0, math.cos(ph), math.sin(ph), math.cos(ph), math.sin(ph), 1.0, 0.0)
            for q in range(n):
                c.add("ApplyRotationX", q, 2.0 * float(bl))
            if shift[0] == "qubit" and shift[1] == l:
                c.add("ApplyRotationX", shift[2], 2.0 * shift[3])
        c.add("QaoaExpectation")

    for l in range(p):
        for e, (_, _, w) in enumerate(prob.edges):
            d = math.pi / (2.0 * w)
            evaluation(("edge", l, e, d))
            evaluation(("edge", l, e, -d))
            plan.append(("gamma", l, 0.5 * w))
        for q in range(n):
            evaluation(("qubit", l, q, math.pi / 4))
            evaluation(("qubit", l, q, -math.pi / 4))
            plan.append(("beta", l, 1.0))
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    with tempfile.TemporaryDirectory() as d:
        cf, rf = os.path.join(d, "c.txt"), os.path.join(d, "r.txt")
        c.write(cf)
        t0 = time.perf_counter()
        r = subprocess.run([drv, cf, "--results", rf], capture_output=True, text=True, timeout=900,
                           env=dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="spread",
                                    OMP_PLACES="cores"))
        t_cpu = time.perf_counter() - t0
        if r.returncode != 0:
            return None
        vals = [float(l.split()[2]) for l in open(rf) if l.startswith("R QaoaExpectation")]
    grad = {("gamma", l): 0.0 for l in range(p)}
    grad.update({("beta", l): 0.0 for l in range(p)})
    for i, (kind, l, coef) in enumerate(plan):
        grad[(kind, l)] += coef * (vals[2 * i] - vals[2 * i + 1])
    diff = max(max(abs(grad[("gamma", l)] - dg[l]), abs(grad[("beta", l)] - db[l])) for l in range(p))
    return {"workload": f"QAOA MaxCut {n} qubits, p = {p}, {len(prob.edges)} edges, complex128",
            "evaluations": len(vals), "cpu_param_shift_s": t_cpu, "cpu_threads": threads,
            "gpu_adjoint_s": t_adj, "cpu_param_shift_over_gpu_adjoint": t_cpu / t_adj,
            "max_abs_diff_gpu_adjoint_vs_cpu_param_shift": diff}


def hea_sample_on_reference(torch, count: int = 24, max_qubits: int = 20, budget_s: float = 30.0):
    """The batched-HEA workload on the host: the first instances of the 512-instance batch
    (seed 20240606) with at most max_qubits qubits, forward pass and <sum_q Z_q>, through the
    compiled reference one instance after another (OpenMP inside each gate), against the
    same instances' costs from the GPU batch engine in complex128."""
    import numpy as np
    from iqs_amd.batch import HeaBatch, instance_circuit, param_offsets, random_batch

    drv = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    if not os.path.exists(drv):
        return None
    nq, dp, ang = random_batch()
    off = param_offsets(nq, dp)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    picked, ref_cost, cpu_s = [], [], 0.0
    with tempfile.TemporaryDirectory() as d:
        for i in range(len(nq)):
            if len(picked) >= count or cpu_s > budget_s:
                break
            if nq[i] > max_qubits:
                continue
            c = instance_circuit(nq[i], dp[i], ang[off[i]:off[i + 1]])
            cf, rf = os.path.join(d, "c.txt"), os.path.join(d, "r.txt")
            c.write(cf)
            r = subprocess.run([drv, cf, "--results", rf, "--time"], capture_output=True, text=True, timeout=600,
                               env=dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="spread",
                                        OMP_PLACES="cores"))
            if r.returncode != 0:
                return None
            cpu_s += sum(float(l.split()[3]) for l in r.stderr.splitlines() if l.startswith("T "))
            z = [float(x) for l in open(rf) if l.startswith("R ZAll") for x in l.split()[2:]]
            picked.append(i)
            ref_cost.append(sum(z))
    b = HeaBatch([nq[i] for i in picked], [dp[i] for i in picked], "c128")
    sample_angles = np.concatenate([ang[off[i]:off[i + 1]] for i in picked])
    gpu_cost = b.evaluate(sample_angles, grad=False)
    b.close()
    return {"workload": f"first {len(picked)} instances with <= {max_qubits} qubits of the 512-instance HEA batch "
                        f"(seed 20240606), forward pass and <sum Z>",
            "instances": len(picked), "qubits": [int(min(nq[i] for i in picked)), int(max(nq[i] for i in picked))],
            "cpu_gate_s": cpu_s, "cpu_threads": threads, "cpu_instances_per_s": len(picked) / cpu_s,
            "max_abs_cost_diff_gpu_c128_vs_reference": float(np.max(np.abs(np.array(gpu_cost) - np.array(ref_cost))))}


def state_error(a, b):
    """(1 - |<b|a>|^2 / (<a|a><b|b>), max |a_i - b_i|) of GPU state a against the complex128
    GPU state b of the same size; a complex64 state is widened into a scratch register."""
    from iqs_amd._lib import call
    from iqs_amd.engine import StateVector
    import numpy as np
    tmp = None
    if a.np_dtype != np.complex128:
        tmp = StateVector(a.num_qubits, "c128")
        call("iqsh_convert", tmp.handle, a.handle)
        a = tmp
    try:
        ov = a.overlap(b)
        f = abs(ov) ** 2 / (a.norm2() * b.norm2())
        return 1.0 - f, a.maxabsdiff(b)
    finally:
        if tmp is not None:
            tmp.close()


def qaoa_measure(n: int, p: int, full: bool, torch):
    """28-qubit QAOA MaxCut gradient: adjoint vs parameter-shift on the same GPU."""
    import numpy as np
    from iqs_amd.qaoa import default_problem

    prob, g, b = default_problem(n, p, "c128")
    prob.adjoint(g, b)  # warm-up (allocations)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    prob.forward_state(g, b)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    cost, dg, db = prob.adjoint(g, b)
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    which = None if full else [("gamma", 0), ("beta", 0)]
    t0 = time.perf_counter()
    ps, nevals = prob.parameter_shift(g, b, which)
    torch.cuda.synchronize()
    t_ps = time.perf_counter() - t0
    total_evals = p * 2 * (len(prob.edges) + n)
    t_ps_full = t_ps if full else t_ps / nevals * total_evals
    diff = max(abs((dg if k == "gamma" else db)[l] - v) for (k, l), v in ps.items())
    fd = prob.finite_difference(g, b, "gamma", p - 1)
    p64, g64, b64 = default_problem(n, p, "c64")
    p64.renorm_every = 2
    p64.adjoint(g64, b64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    c2, dg2, db2 = p64.adjoint(g64, b64)
    torch.cuda.synchronize()
    t_adj64 = time.perf_counter() - t0
    # final complex64 state against the complex128 one (FP64 reductions on the GPU)
    inf64, mx64 = state_error(p64.forward_state(g64, b64), prob.forward_state(g, b))
    return {
        "qubits": n, "layers": p, "edges": len(prob.edges), "params": 2 * p, "dtype": "f64",
        "cost": cost, "forward_s": t_fwd, "adjoint_grad_s": t_adj,
        "param_shift_evals_total": total_evals, "param_shift_evals_measured": nevals,
        "param_shift_s_measured": t_ps,
        "param_shift_grad_s": t_ps_full, "param_shift_full_run": full,
        "speedup_adjoint_vs_param_shift": t_ps_full / t_adj,
        "max_abs_diff_adjoint_vs_param_shift": diff,
        "central_fd_h1e-4_gamma_last": fd, "adjoint_gamma_last": dg[-1],
        "c64": {"adjoint_grad_s": t_adj64, "cost_abs_err_vs_f64": abs(c2 - cost),
                "grad_max_abs_err_vs_f64": float(max(np.abs(dg2 - dg).max(), np.abs(db2 - db).max())),
                "state_one_minus_fidelity_vs_f64": inf64, "state_max_abs_err_vs_f64": mx64},
    }


def brick_fused_measure(sv, layers, torch):
    """The timed brick circuit again through the drop-in with fusion on (TurnOnFusion,
    <= 5-position blocks on the matrix cores): gates/s of the whole queued-and-flushed run
    and its distance from the unfused state `sv` (complex128, same gates)."""
    import ctypes as C

    import intelqs_py as iqs
    from iqs_amd import runner
    from iqs_amd._lib import call
    from iqs_amd.circuits import Circuit

    n = sv.num_qubits
    c = Circuit(n)
    for ops in layers:
        for op in ops:
            c.add(*op)
    reg = iqs.QubitRegister(n, "base", 0, 0)
    reg.SetFusionBlockQubits(5)
    reg.TurnOnFusion()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    runner.run(c, reg)  # ends with TurnOffFusion: every queued block applied
    reg.Sync()
    t = time.perf_counter() - t0
    gates, blocks = reg.FusionStats()
    mx = C.c_double()
    call("iqsh_maxabsdiff", C.c_void_p(reg.DeviceHandle()), sv.handle, 1.0, 0.0, C.byref(mx))
    del reg
    return {"workload": f"the timed {n}-qubit brick layers ({len(c.ops)} gates) through the drop-in with "
                        f"TurnOnFusion, <= 5-position blocks; not the headline value (that is one pass per gate)",
            "gates": gates, "blocks": blocks, "s": t, "gates_per_s": gates / t,
            "max_abs_diff_vs_unfused": mx.value}


def fusion_measure(n: int, ngates: int, torch):
    """Random Haar 1q/2q circuit (seed 20240603) fused into <= 5-qubit blocks through the
    drop-in (complex64 and complex128), against the FP64 unfused run of the same gates."""
    import ctypes as C

    import intelqs_py as iqs
    from iqs_amd import runner
    from iqs_amd._lib import call
    from iqs_amd.circuits import random_1q2q_circuit

    def handle(reg):
        return C.c_void_p(reg.DeviceHandle())

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return out, time.perf_counter() - t0

    def compare(reg, ref):
        """(1 - fidelity, max-abs error) of reg against ref (complex128), FP64 reductions."""
        dre, dim, n2a, n2b, mx = (C.c_double() for _ in range(5))
        a = handle(reg)
        tmp = None
        if isinstance(reg, iqs.QubitRegisterSP):
            tmp = C.c_void_p()
            call("iqsh_create", n, 1, C.byref(tmp))
            call("iqsh_convert", tmp, a)
            a = tmp
        b = handle(ref)
        call("iqsh_overlap", a, b, C.byref(dre), C.byref(dim))
        call("iqsh_norm2", a, C.byref(n2a))
        call("iqsh_norm2", b, C.byref(n2b))
        call("iqsh_maxabsdiff", a, b, 1.0, 0.0, C.byref(mx))
        if tmp is not None:
            call("iqsh_destroy", tmp)
        f = (dre.value ** 2 + dim.value ** 2) / (n2a.value * n2b.value)
        return 1.0 - f, mx.value, n2a.value

    c = random_1q2q_circuit(n, ngates, 20240603)
    ref = iqs.QubitRegister(n, "base", 0, 0)
    _, t_ref = timed(lambda: runner.run(c, ref))
    out = {"qubits": n, "gates": ngates, "block_qubits": 5, "seed": 20240603,
           "f64_unfused_s": t_ref, "f64_unfused_gates_per_s": ngates / t_ref}
    for prec, cls in (("c128", iqs.QubitRegister), ("c64", iqs.QubitRegisterSP)):
        c.precision = "sp" if prec == "c64" else "dp"
        if prec == "c64":
            reg = cls(n, "base", 0, 0)
            _, t_unf = timed(lambda: runner.run(c, reg))
            inf_u, mx_u, _ = compare(reg, ref)
            del reg
        reg = cls(n, "base", 0, 0)
        reg.SetFusionBlockQubits(5)
        reg.TurnOnFusion()
        _, t_f = timed(lambda: runner.run(c, reg))
        gates, blocks = reg.FusionStats()
        inf, mx, n2 = compare(reg, ref)
        S = (1 << n) * (16 if prec == "c128" else 8)
        o = {"fused_s": t_f, "fused_gates_per_s": ngates / t_f, "blocks": blocks, "gates_fused": gates,
             "block_GBs_avg": 2 * S * blocks / t_f / 1e9, "one_minus_fidelity": inf, "max_abs_err": mx,
             "norm2": n2}
        if prec == "c64":
            o.update({"unfused_s": t_unf, "unfused_one_minus_fidelity": inf_u, "unfused_max_abs_err": mx_u})
        out[prec] = o
        del reg
    c.precision = "dp"
    return out


def batch_measure(reps: int, torch, tune=None):
    """512 hardware-efficient-ansatz instances (n ~ U{10..22}, d ~ U{10..50}, seed 20240606):
    forward + adjoint gradient of every angle in one persistent, graph-replayed launch.
    complex64 (autotuned) timed; complex128 run of the same batch as the FP64 reference;
    parameter-shift spot check (one angle per instance) of the complex128 gradient."""
    import numpy as np
    from iqs_amd._lib import call
    from iqs_amd.batch import HeaBatch, random_batch

    nq, dp, ang = random_batch()
    out = {"instances": len(nq), "seed": 20240606, "qubits": [min(nq), max(nq)], "depths": [min(dp), max(dp)],
           "params": int(2 * sum(n * d for n, d in zip(nq, dp)))}
    b64 = HeaBatch(nq, dp, "c64")
    b64.set_angles(ang)
    if tune is None:
        log("HEA batch: autotuning complex64")
        t0 = time.perf_counter()
        tune, best_ms, trials = b64.autotune(grad=True, reps=2)
        log(f"HEA batch: tuned {tune} at {best_ms:.2f} ms")
        out["autotune_s"] = time.perf_counter() - t0
    else:  # a given tuning (profiling runs): no autotune trials
        b64.tune = tune
        tune, trials = b64.tune, []
    out["tune"] = {"variant": tune[0], "workers_per_cu": tune[1], "order": tune[2], "group_mib": tune[3],
                   "tile": HeaBatch.variants("c64")[tune[0]]}
    out["trials_ms"] = [[list(t), round(ms, 3)] for t, ms in trials]

    def timed(fn, k):
        fn()
        call("iqsh_sync")
        call("iqsh_timer_start")
        for _ in range(k):
            fn()
        ms = C.c_double()
        call("iqsh_timer_stop", C.byref(ms))
        return ms.value / k

    ms_grad = timed(lambda: b64.run(grad=True), reps)
    ms_fwd = timed(lambda: b64.run(grad=False), reps)
    c64, g64 = b64.results()
    traffic = b64.traffic(grad=True)
    out["c64"] = {"ms_forward_and_gradient": ms_grad, "ms_forward_only": ms_fwd,
                  "instances_per_s": len(nq) / (ms_grad * 1e-3),
                  "algorithmic_bytes": traffic, "algorithmic_GBs": traffic / (ms_grad * 1e-3) / 1e9}
    # Bound of the persistent kernel, from the committed counter summary: VALU issue (busy
    # fraction of the tuned complex64 gradient run) and its measured HBM traffic.
    busy = profile_value("k_hea_batch512_pmc.md", "valu_busy_c64_grad")
    hbm = profile_value("k_hea_batch512_pmc.md", "traffic_bytes_c64_grad")
    out["c64"]["roofline"] = {"bound": "valu", "valu_busy_frac": busy,
                              "hbm_traffic_bytes": hbm,
                              "hbm_frac": hbm / (ms_grad * 1e-3) / (HBM_PEAK_GBS * 1e9) if hbm else None,
                              "source": "profiles/k_hea_batch512_pmc.md (rocprofv3 counters of the same tuned run)"}
    log("HEA batch: complex128 reference run")
    b128 = HeaBatch(nq, dp, "c128")
    c128, g128 = b128.evaluate(ang)
    out["c128"] = {"ms_forward_and_gradient": timed(lambda: b128.run(grad=True), max(1, reps // 2))}
    ps = b128.parameter_shift(ang, [0] * len(nq))
    # final states of every instance, complex64 against complex128
    b64.run(grad=False, keep_state=True)
    b128.set_angles(ang)
    b128.run(grad=False, keep_state=True)
    inf, mx = [], []
    for i in range(len(nq)):
        a, r = b64.state(i).astype(np.complex128), b128.state(i)
        f = abs(np.vdot(r, a)) ** 2 / (np.vdot(a, a).real * np.vdot(r, r).real)
        inf.append(1.0 - f)
        mx.append(float(np.abs(a - r).max()))
    out["c64"].update({"state_one_minus_fidelity_vs_f64_max": float(max(inf)),
                       "state_one_minus_fidelity_vs_f64_median": float(np.median(inf)),
                       "state_max_abs_err_vs_f64": float(max(mx))})
    b64.close()
    b128.close()
    out["c128"]["grad_vs_param_shift_max_abs"] = float(np.max(np.abs(g128[np.cumsum([0] + [2 * n * d for n, d in zip(nq, dp)])[:-1]] - ps)))
    out["c64"].update({"cost_max_abs_err_vs_f64": float(np.max(np.abs(c64 - c128))),
                       "grad_max_abs_err_vs_f64": float(np.max(np.abs(g64 - g128))),
                       "grad_rms_err_vs_f64": float(np.sqrt(np.mean((g64 - g128) ** 2)))})
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--sweep-reps", type=int, default=3)
    ap.add_argument("--sweep-c64-qubits", type=int, default=32, help="complex64 sweep register (0 = none)")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate", action="store_true", help="one 2^28-amplitude complex128 copy, for PMC scaling")
    ap.add_argument("--no-qaoa", action="store_true")
    ap.add_argument("--qaoa-sample", action="store_true",
                    help="parameter-shift for gamma_1 and beta_1 only (the full run takes about 100 s)")
    ap.add_argument("--qaoa-qubits", type=int, default=28)
    ap.add_argument("--no-fusion", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the HEA batch section (for profiling its kernels); no metric value")
    ap.add_argument("--batch-reps", type=int, default=5)
    ap.add_argument("--batch-tune", default="",
                    help="variant,workers_per_cu,order,group_mib: skip the autotuner (profiling runs)")
    ap.add_argument("--fusion-qubits", type=int, default=32)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: --gpus {args.gpus} is not supported: this engine keeps the whole register in one "
              f"GPU's HBM and runs on one GPU only (use --gpus 1).", file=sys.stderr)
        return 2

    if args.calibrate:
        return calibrate(28)

    import numpy as np
    import torch

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        return 1
    from iqs_amd import gates as G
    from iqs_amd._lib import call
    from iqs_amd.circuits import SplitMix64, brick_layer_ops
    from iqs_amd.engine import StateVector

    torch.cuda.init()
    stream = torch.cuda.current_stream()
    call("iqsh_set_stream", stream.cuda_stream)

    batch_tune = tuple(int(x) for x in args.batch_tune.split(",")) if args.batch_tune else None
    if args.batch_only:
        print(json.dumps({"hea_batch": batch_measure(args.batch_reps, torch, batch_tune)}))
        return 0

    n = args.qubits
    S = (1 << n) * 16
    sv = StateVector(n, "c128")

    rng = SplitMix64(20240602)
    layers = [brick_layer_ops(n, l, rng) for l in range(args.steps)]
    wrng = SplitMix64(1)
    warm = [brick_layer_ops(n, l, wrng) for l in range(args.warmup)]

    sv.init_basis(0)
    for ops in warm:
        for op in ops:
            apply_op(sv, op, G)
    sv.init_basis(0)
    torch.cuda.synchronize()
    log(f"{n}-qubit register ready, {args.warmup} warm-up layers done")

    # ---- timed region: K layers, an event pair around every gate ----
    ev = []
    t0 = time.perf_counter()
    for ops in layers:
        for op in ops:
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            apply_op(sv, op, G)
            b.record(stream)
            ev.append((kernel_kind(op), a, b))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ngates = len(ev)
    per_kind = {}
    for kind, a, b in ev:
        per_kind.setdefault(kind, []).append(a.elapsed_time(b) * 1e-3)
    dense = per_kind["dense1q_pair"]
    dense_avg = sum(dense) / len(dense)
    achieved = 2 * S / dense_avg / 1e9

    readout = [1.0 - 2.0 * sv.prob1(q) for q in range(n)]
    log(f"timed brick circuit: {ngates} gates in {elapsed:.2f} s")
    brick_fused = None if args.no_fusion else brick_fused_measure(sv, layers, torch)
    if brick_fused:
        log(f"fused brick circuit: {brick_fused['gates']} gates in {brick_fused['blocks']} blocks, {brick_fused['s']:.2f} s")

    # ---- the same circuit undone: distance back to |0...0> (FP64 error of the round trip) ----
    for ops in reversed(layers):
        for op in reversed(ops):
            apply_inverse_op(sv, op, G)
    ref0 = StateVector(n, "c128")
    ref0.init_basis(0)
    roundtrip = sv.maxabsdiff(ref0)
    ref0.close()
    log(f"brick circuit undone: max |psi - |0>| = {roundtrip:.3e}")

    # ---- attainable streaming rate on this GPU: the library's in-place scale of the whole
    # state (k_scale, iqsh_scale: one read and one write per amplitude, no partner), the
    # simplest read-modify-write pass; the dense 1q update is compared against it ----
    stream_ms = _time_ms(lambda: sv.scale(-1.0), 5, stream, torch)
    attainable = 2 * S / stream_ms / 1e6
    log(f"in-place scale of the state: {attainable:.0f} GB/s")

    # ---- per-kind, permutation and reduction sweeps over every position ----
    sweep = sweep_c64 = perm = red = None
    if not args.no_sweep:
        sweep = kind_sweep(sv, n, 16, args.sweep_reps, stream, torch)
        perm = permute_sweep(sv, n, 16, args.sweep_reps, stream, torch)
        red = reduction_sweep(sv, n, 16, args.sweep_reps, stream, torch)
    del sv
    if not args.no_sweep and args.sweep_c64_qubits:
        n64 = args.sweep_c64_qubits
        sv64 = StateVector(n64, "c64")
        sv64.init_uniform(1.0 / math.sqrt(float(1 << n64)))
        sweep_c64 = kind_sweep(sv64, n64, 8, args.sweep_reps, stream, torch)
        sweep_c64["permute"] = permute_sweep(sv64, n64, 8, args.sweep_reps, stream, torch)
        sweep_c64["reductions"] = reduction_sweep(sv64, n64, 8, args.sweep_reps, stream, torch)
        sweep_c64["qubits"] = n64
        sv64.close()
    pool = None if args.no_sweep else pool_sweep(args.sweep_reps, stream, torch)
    log("brick circuit and per-kind sweep done")
    qaoa = None if args.no_qaoa else qaoa_measure(args.qaoa_qubits, 8, not args.qaoa_sample, torch)
    log("QAOA gradient done")
    fusion = None if args.no_fusion else fusion_measure(args.fusion_qubits, 200, torch)
    log("fusion done")
    batch = None if args.no_batch else batch_measure(args.batch_reps, torch, batch_tune)
    log("HEA batch done")
    cpu = None if args.no_cpu else cpu_baseline(n, layers[0])
    if cpu is not None:
        cpu["per_kind_30q"] = cpu_kind_sample(n)
        cpu["brick_20q_depth20"] = brick20_against_reference(torch)
        cpu["qaoa_16q_param_shift"] = None if args.no_qaoa else qaoa_param_shift_on_reference(torch)
        cpu["hea_batch_sample"] = None if args.no_batch else hea_sample_on_reference(torch)
    log("CPU baseline done")

    out = {
        "metric": METRIC,
        "value": ngates / elapsed,
        "unit": "gates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{n}-qubit ComplexDP random brick circuit, depth {args.steps} (H, R{{X,Y,Z}}, CNOT ladder "
                               f"per layer), seed 20240602, state resident in HBM",
                   "qubits": n, "depth": args.steps, "gates": ngates},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profile(n),
                     "kernel": "k_pair<double,1,false,true,128,2,4,SL> (dense 1q update, target >= 6: H, RX, RY; "
                               "sub-ranges 2^SL items apart, SL=10 for targets 6-10, 11 above)",
                     "avg_launch_ms": 1e3 * dense_avg, "launches": len(dense),
                     "attainable": attainable, "frac_of_attainable": achieved / attainable,
                     "attainable_kernel": "k_scale (iqsh_scale on the same state: in-place read and write of "
                                          "every amplitude, measured in this run)"},
        "per_kind_avg_ms": {k: 1e3 * sum(v) / len(v) for k, v in per_kind.items()},
        "hbm_frac_1q": achieved / HBM_PEAK_GBS,
        "hbm_frac_1q_all_targets": (2 * S / (sum(dense + per_kind.get("dense1q_lane", [])) /
                                             len(dense + per_kind.get("dense1q_lane", []))) / 1e9) / HBM_PEAK_GBS,
        "z_readout_first4": readout[:4],
        "brick_roundtrip_max_abs_err": roundtrip,
        "brick_fused": brick_fused,
        "sweep": sweep,
        "sweep_permute": perm,
        "sweep_reductions": red,
        "sweep_c64": sweep_c64,
        "sweep_pool": pool,
        "qaoa_grad": qaoa,
        "fusion": fusion,
        "hea_batch": batch,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
