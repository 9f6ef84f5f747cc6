"""Noise gates of the drop-in (SetNoiseTimescales / ApplyNoiseGate, src/qureg_noisysimul.cpp)
against the compiled reference on the same seeded generator: the Gaussian Pauli-twirl
angles come from the register's mt19937 state stream, so the final state must match the
reference's (noise_dp_state.npz) to 1e-12 and its read-outs to 1e-12."""
import os

import numpy as np
import pytest

from iqs_amd import runner
from iqs_amd.circuits import Circuit

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_noise_gates_match_reference():
    c = Circuit.parse(open(os.path.join(GOLDEN, "noise_dp.noisecirc")).read())
    reg, out = runner.run(c)
    want = np.load(os.path.join(GOLDEN, "noise_dp_state.npz"))["state"]
    got = np.array(reg, copy=True)
    assert np.max(np.abs(got - want)) < 1e-12
    ref = [l for l in open(os.path.join(GOLDEN, "noise_dp.noiseres")).read().splitlines() if l]
    assert len(out) == len(ref)
    for a, b in zip(out, ref):
        ta, tb = a.split(), b.split()
        assert ta[:2] == tb[:2]
        assert np.allclose([float(x) for x in ta[2:]], [float(x) for x in tb[2:]], rtol=0, atol=1e-12)
    assert abs(reg.GetT1() - 50.0) < 1e-15 and abs(reg.GetT2() - 30.0) < 1e-15
    assert abs(reg.GetTphi() - 1.0 / (1.0 / 30.0 - 1.0 / 100.0)) < 1e-12


def test_noise_gate_needs_generator_and_valid_timescales():
    c = Circuit(3)
    reg = runner.new_register(c)
    with pytest.raises(ValueError):
        reg.SetNoiseTimescales(10.0, 4.0)  # T2 < T1 / 2
    reg.SetNoiseTimescales(10.0, 8.0)
    with pytest.raises(RuntimeError):
        reg.ApplyNoiseGate(0, 1.0)  # no generator attached


NOISE_ONLY = ("SetGateDurations", "SetDecoherenceTime", "ResetTimeForAllQubits", "ApplyNoiseGatesOnAllQubits", "Counts")


def _run_noisy(text, clean=False):
    """Execute a NoisyQureg program (tests/golden/make_golden.py:noisy_program) on the GPU
    register; returns the register and its "R ..." read-out lines.  clean=True runs the same
    logical gates on a plain QubitRegister."""
    import intelqs_py as m
    psi, out = None, []
    for line in text.splitlines():
        t = line.split()
        if not t or t[0].startswith("#"):
            continue
        op, a = t[0], t[1:]
        if clean and op in NOISE_ONLY:
            continue
        if op == "NoisyQureg":
            psi = (m.QubitRegister(int(a[0]), "base", 0, 0) if clean
                   else m.NoisyQureg(int(a[0]), int(a[1]), float(a[2]), float(a[3])))
        elif op == "Initialize":
            psi.Initialize(a[0], int(a[1]))
        elif op in ("SetGateDurations", "SetDecoherenceTime"):
            getattr(psi, op)(float(a[0]), float(a[1]))
        elif op in ("ResetTimeForAllQubits", "ApplyNoiseGatesOnAllQubits"):
            getattr(psi, op)()
        elif op == "ApplyHadamard":
            psi.ApplyHadamard(int(a[0]))
        elif op.startswith("ApplyRotation"):
            getattr(psi, op)(int(a[0]), float(a[1]))
        elif op == "ApplyCPauliX":
            psi.ApplyCPauliX(int(a[0]), int(a[1]))
        elif op in ("Apply1QubitGate", "ApplyControlled1QubitGate"):
            nq = 1 if op == "Apply1QubitGate" else 2
            v = [float(x) for x in a[nq:]]
            mat = np.array([complex(v[i], v[i + 1]) for i in range(0, 8, 2)]).reshape(2, 2)
            getattr(psi, op)(*[int(x) for x in a[:nq]], mat)
        elif op == "Counts":
            out.append(f"R Counts {psi.GetTotalExperimentalGateCount()} {psi.GetOneQubitExperimentalGateCount()} "
                       f"{psi.GetTwoQubitExperimentalGateCount()}")
            for r in range(psi.NumQubits()):
                out.append(f"R Row {r} " + " ".join(str(c) for c in psi.GetExperimentalGateCount(r)))
        else:
            raise ValueError(op)
    return psi, out


def test_noisy_qureg_matches_reference():
    """Two trajectories of iqs::NoisyQureg (include/NoisyQureg.hpp of the reference, compiled
    by oracle/noisy_driver.cpp): the same default_random_engine + normal_distribution stream
    drives the noise rotations, so the final state matches the reference's to 1e-12 and the
    experimental-gate counts exactly."""
    text = open(os.path.join(GOLDEN, "noisy_dp.nqprog")).read()
    psi, out = _run_noisy(text)
    want = np.load(os.path.join(GOLDEN, "noisy_dp_state.npz"))["state"]
    got = np.array(psi, copy=True)
    assert np.max(np.abs(got - want)) < 1e-12
    ref = [l for l in open(os.path.join(GOLDEN, "noisy_dp.nqres")).read().splitlines() if l]
    assert out == ref
    # The noise changed the state: the noiseless circuit gives a different one.
    clean, _ = _run_noisy(text, clean=True)
    assert np.max(np.abs(np.array(clean, copy=True) - want)) > 1e-3


def test_noisy_qureg_invalid_times():
    import intelqs_py as m
    psi = m.NoisyQureg(3, 7, -5.0, 10.0)
    psi.ApplyHadamard(0)  # idle time 0: no noise gate yet
    with pytest.raises(ValueError):
        psi.ApplyHadamard(1)  # qubit 1 idled one gate with T1 < 0
