"""Pool of states (intelqs_py.StatePool over iqsh_pool_*): the single-GPU form of the
reference's pool, one state per MPI rank group (src/mpi_env.cpp:239-404).

  * Trajectories: the noisy-pool tutorial's program (tutorials/get_started_with_noisy_IQS
    .cpp) on 5 states; state s must equal the compiled reference run with seed + s
    (tests/golden/pool_dp_*: same state stream as the reference's pool state s) to 1e-12,
    read-outs per state to 1e-12, the incoherent sum to their sum.
  * Common gates on Register() act on every state as on a lone register (the oracle runs
    the same gate list), padding states stay zero.
  * Per-state reductions equal numpy on the downloaded states; per-state 1- and 2-qubit
    channel branches equal ApplyChannel on a lone register seeded as that pool state."""
import os

import numpy as np
import pytest

import oracle
from iqs_amd.circuits import Circuit

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _m():
    import intelqs_py
    return intelqs_py


def _run_pool_program(pool, c):
    reg = pool.Register()
    out = []
    for op in c.ops:
        name, args = op[0], op[1:]
        if name == "SetRngSeed":
            continue  # the pool seeds every state at once
        if name in ("SetNoiseTimescales", "ApplyNoiseGate"):
            getattr(pool, name)(*args)
        elif name == "GetProbability":
            out.append((args[0], pool.GetProbability(args[0])))
        else:
            getattr(reg, name)(*args)
    return out


def test_pool_trajectories_match_reference():
    c = Circuit.parse(open(os.path.join(GOLDEN, "pool_dp.poolcirc")).read())
    ref = np.load(os.path.join(GOLDEN, "pool_dp_states.npz"))
    states, seed = ref["states"], int(ref["seed"])
    m = _m()
    pool = m.StatePool(c.num_qubits, len(states))
    pool.SetSeedStreamPtrs(seed)
    out = _run_pool_program(pool, c)
    for s in range(len(states)):
        assert np.max(np.abs(pool.GetState(s) - states[s])) < 1e-12, s
    want = {}
    for block in open(os.path.join(GOLDEN, "pool_dp.poolres")).read().split("# state ")[1:]:
        lines = block.splitlines()
        s = int(lines[0])
        for l in lines[1:]:
            t = l.split()
            want.setdefault(int(t[2]), [None] * len(states))[s] = float(t[3])
    for q, probs in out:
        assert np.max(np.abs(np.array(probs) - np.array(want[q]))) < 1e-12
        assert abs(m.StatePool.IncoherentSumOverAllStatesOfPool(probs) - sum(want[q])) < 1e-12
    # Padding states (5 of 8 slots used) are untouched.
    full = np.array(pool.Register(), copy=True)
    assert np.all(full[len(states) << c.num_qubits:] == 0)


def test_common_gates_act_on_every_state():
    m = _m()
    n, k = 7, 3
    pool = m.StatePool(n, k, "++++", 0)
    c = Circuit(n, precision="dp", init="plus")
    rs = np.random.default_rng(17)
    for _ in range(30):
        q = int(rs.integers(n))
        kind = int(rs.integers(4))
        if kind == 0:
            c.add("ApplyRotationX", q, float(rs.uniform(0, 6)))
        elif kind == 1:
            c.add("ApplyHadamard", q)
        elif kind == 2:
            c.add("ApplyCPauliX", q, (q + 1 + int(rs.integers(n - 1))) % n)
        else:
            c.add("ApplySwap", q, (q + 2) % n)
    reg = pool.Register()
    for op in c.ops:
        getattr(reg, op[0])(*op[1:])
    want, _ = oracle.run(c.to_text())
    for s in range(k):
        assert np.max(np.abs(pool.GetState(s) - want)) < 1e-12, s
    assert np.allclose(pool.ComputeNorm(), 1.0, atol=1e-12)


def test_per_state_reductions_and_overlap():
    c = Circuit.parse(open(os.path.join(GOLDEN, "pool_dp.poolcirc")).read())
    m = _m()
    ns = 6
    pool = m.StatePool(c.num_qubits, ns)
    pool.SetSeedStreamPtrs(99)
    _run_pool_program(pool, c)
    pool.Register().ApplyHadamard(2)
    st = [pool.GetState(s) for s in range(ns)]
    idx = np.arange(st[0].size)
    for q in (0, 4):
        want = [np.sum(np.abs(v[(idx >> q) & 1 == 1]) ** 2) for v in st]
        assert np.max(np.abs(np.array(pool.GetProbability(q)) - want)) < 1e-13
    mask = (1 << 1) | (1 << 5)
    sign = np.where(np.array([bin(i & mask).count("1") & 1 for i in idx]) == 1, -1.0, 1.0)
    want = [np.sum(sign * np.abs(v) ** 2) for v in st]
    assert np.max(np.abs(np.array(pool.ExpectationValueZString([1, 5])) - want)) < 1e-13
    assert np.max(np.abs(np.array(pool.ComputeNorm()) - [np.linalg.norm(v) for v in st])) < 1e-13
    ideal = m.QubitRegister(c.num_qubits, "base", 0, 0)
    ideal.ApplyHadamard(0)
    ideal.ApplyRotationY(3, 0.7)
    iv = np.array(ideal, copy=True)
    got = pool.ComputeOverlap(ideal)
    want = [np.vdot(iv, v) for v in st]
    assert np.max(np.abs(np.array(got) - want)) < 1e-13


def test_per_state_channel_matches_lone_registers():
    m = _m()
    n, ns, seed = 5, 4, 31
    chi = np.diag([0.6, 0.25, 0.1, 0.05]).astype(complex)
    cm = m.CM4x4(chi)
    cm.SolveEigenSystem()
    pool = m.StatePool(n, ns, "++++", 0)
    pool.SetSeedStreamPtrs(seed)
    for step in range(6):
        pool.ApplyChannel(step % n, cm)
        pool.Register().ApplyCPauliX(step % n, (step + 2) % n)
    for s in range(ns):
        reg = m.QubitRegister(n, "++++", 0, 0)
        rng = m.RandomNumberGenerator()
        rng.SetSeedStreamPtrs(seed + s)  # its state stream is pool state s's
        reg.SetRngPtr(rng)
        for step in range(6):
            reg.ApplyChannel(step % n, cm)
            reg.ApplyCPauliX(step % n, (step + 2) % n)
        reg.ResetRngPtr()
        assert np.max(np.abs(pool.GetState(s) - np.array(reg, copy=True))) < 1e-12, s


@pytest.mark.parametrize("order", ["high_first", "low_first"])
def test_per_state_2q_channel_matches_lone_registers(order):
    """2-qubit channel branches (a dense complex chi, so every Pauli pair term is present)
    drawn per state; both qubit orders pin the 2 b(qubit1) + b(qubit2) row convention of
    the per-state kernel against Apply2QubitGate.  Three states: the padding state stays 0."""
    m = _m()
    n, ns, seed = 6, 3, 57
    r = np.random.default_rng(12)
    a = r.normal(size=(16, 16)) + 1j * r.normal(size=(16, 16))
    chi = a @ a.conj().T
    chi /= np.trace(chi).real
    cm = m.CM16x16(chi)
    cm.SolveEigenSystem()
    pairs = [(4, 1), (0, 5), (3, 2)] if order == "high_first" else [(1, 4), (5, 0), (2, 3)]
    pool = m.StatePool(n, ns, "++++", 0)
    pool.SetSeedStreamPtrs(seed)
    for step in range(5):
        q1, q2 = pairs[step % 3]
        pool.ApplyChannel(q1, q2, cm)
        pool.Register().ApplyRotationY(step % n, 0.3 + step)
    signs = pool.GetOverallSignOfChannels()
    for s in range(ns):
        reg = m.QubitRegister(n, "++++", 0, 0)
        rng = m.RandomNumberGenerator()
        rng.SetSeedStreamPtrs(seed + s)
        reg.SetRngPtr(rng)
        for step in range(5):
            q1, q2 = pairs[step % 3]
            reg.ApplyChannel(q1, q2, cm)
            reg.ApplyRotationY(step % n, 0.3 + step)
        reg.ResetRngPtr()
        want = np.array(reg, copy=True)
        assert np.max(np.abs(want)) > 1e-3
        assert np.max(np.abs(pool.GetState(s) - want)) < 1e-12, s
        assert signs[s] == reg.GetOverallSignOfChannels()
    full = np.array(pool.Register(), copy=True)
    assert np.all(full[ns << n:] == 0)
    with pytest.raises(ValueError):
        pool.ApplyChannel(2, 2, cm)


def test_pool_errors():
    m = _m()
    pool = m.StatePool(4, 3)
    with pytest.raises(IndexError):
        pool.GetProbability(4)
    with pytest.raises(IndexError):
        pool.GetState(3)
    with pytest.raises(ValueError):
        m.StatePool(4, 0)
    with pytest.raises(ValueError):
        pool.Initialize("rand", 0)
    with pytest.raises(ValueError):
        m.MPIEnvironment.UpdateStateComm(4)


@pytest.mark.parametrize("dt", ["c128", "c64"])
def test_per_state_kernels_every_form(dt):
    """iqsh_pool_* on 9-qubit states (3 states in 4 slots, the fourth left as uploaded):
    per-state 1q at positions below 6 (lane-shuffle form) and from 6 up (pair tiles), per-state
    2q with both positions inside a 128-byte line and outside it, then every per-state
    reduction, each against numpy on the same states."""
    import ctypes as C

    from iqs_amd._lib import call
    from iqs_amd.circuits import SplitMix64, haar_unitary
    from iqs_amd.engine import StateVector

    sq, ns, n = 9, 3, 11
    npdt = np.complex128 if dt == "c128" else np.complex64
    tol = 1e-12 if dt == "c128" else 2e-5
    r = np.random.default_rng(61)
    host = (r.normal(size=1 << n) + 1j * r.normal(size=1 << n)).astype(npdt)
    sv = StateVector(n, dt)
    sv.upload(host)
    want = host.astype(np.complex128).reshape(1 << (n - sq), 1 << sq)
    rng = SplitMix64(62)

    def mats(k):
        us = [haar_unitary(1 << k, rng) for _ in range(ns)]
        flat = np.concatenate([np.stack([u.real, u.imag], -1).reshape(-1) for u in us]).astype(np.float64)
        return us, flat

    for pos in (0, 3, 5, 6, 8):
        us, flat = mats(1)
        call("iqsh_pool_apply_1q_each", sv.handle, sq, pos, flat.ctypes.data_as(C.POINTER(C.c_double)), ns)
        for s in range(ns):
            v = want[s].reshape(1 << (sq - 1 - pos), 2, 1 << pos)
            want[s] = np.einsum("ij,ajb->aib", us[s].astype(npdt).astype(np.complex128), v).reshape(-1)
    idx = np.arange(1 << sq)
    for ph, pl in ((1, 0), (2, 7), (8, 6)):
        us, flat = mats(2)
        call("iqsh_pool_apply_2q_each", sv.handle, sq, ph, pl, flat.ctypes.data_as(C.POINTER(C.c_double)), ns)
        base = idx[((idx >> ph) & 1) == 0]
        base = base[((base >> pl) & 1) == 0]
        rows = [base, base | (1 << pl), base | (1 << ph), base | (1 << ph) | (1 << pl)]
        for s in range(ns):
            x = np.stack([want[s][rw] for rw in rows])
            y = us[s].astype(npdt).astype(np.complex128) @ x
            for t, rw in enumerate(rows):
                want[s][rw] = y[t]
    got = sv.download().astype(np.complex128).reshape(1 << (n - sq), 1 << sq)
    assert np.abs(got[:ns] - want[:ns]).max() < tol
    assert np.array_equal(got[ns:], host.astype(np.complex128).reshape(1 << (n - sq), 1 << sq)[ns:])
    out = (C.c_double * (2 * ns))()
    for kind, arg, f in ((0, 0, lambda v: np.sum(np.abs(v) ** 2)),
                         (1, 4, lambda v: np.sum(np.abs(v[(idx >> 4) & 1 == 1]) ** 2)),
                         (2, 0b100010001, lambda v: np.sum(np.where(
                             np.array([bin(i & 0b100010001).count("1") & 1 for i in idx]) == 1, -1.0, 1.0)
                             * np.abs(v) ** 2))):
        call("iqsh_pool_reduce", sv.handle, sq, kind, arg, out, ns)
        assert np.max(np.abs(np.array(out[:ns]) - [f(got[s]) for s in range(ns)])) < 1e-11
    ref_h = (r.normal(size=1 << sq) + 1j * r.normal(size=1 << sq)).astype(npdt)
    ref = StateVector(sq, dt)
    ref.upload(ref_h)
    call("iqsh_pool_overlap", sv.handle, sq, ref.handle, out, ns)
    wov = [np.vdot(ref_h.astype(np.complex128), got[s]) for s in range(ns)]
    assert np.max(np.abs(np.array(out[0::2]) + 1j * np.array(out[1::2]) - wov)) < 1e-11
