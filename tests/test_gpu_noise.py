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
