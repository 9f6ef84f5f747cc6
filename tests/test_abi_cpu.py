"""CPU-side checks of the native boundary (no GPU calls): the C-ABI library loads and
exports every function that include/iqs_hip.h declares; the drop-in library and module
load and expose the reference's method names."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "intel-qs_amd")


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(iqsh_[a-z0-9_]+)\s*\(", text)))


def test_c_abi_exports_every_declared_symbol():
    lib = ctypes.CDLL(os.path.join(PKG, "lib", "libiqs_hip.so"))
    names = declared("iqs_hip.h")
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_dropin_library_loads():
    ctypes.CDLL(os.path.join(PKG, "lib", "libiqs_amd.so"))


REFERENCE_METHODS = [
    # pybind11/intelqs_py.cpp:173-352 of the reference module
    "NumQubits", "GlobalSize", "LocalSize", "ApplyRotationX", "ApplyRotationY", "ApplyRotationZ",
    "ApplyPauliX", "ApplyPauliY", "ApplyPauliZ", "ApplyPauliSqrtX", "ApplyPauliSqrtY", "ApplyPauliSqrtZ",
    "ApplyT", "ApplyRotationXY", "ApplyHadamard", "ApplySwap", "ApplyCRotationX", "ApplyCRotationY",
    "ApplyCRotationZ", "ApplyCPauliX", "ApplyCPauliY", "ApplyCPauliZ", "ApplyCPauliSqrtZ", "ApplyCHadamard",
    "Apply1QubitGate", "ApplyControlled1QubitGate", "ApplyToffoli", "Initialize", "TurnOnSpecialize",
    "TurnOffSpecialize", "TurnOnSpecializeV2", "TurnOffSpecializeV2", "ResetRngPtr", "SetRngPtr",
    "SetSeedRngPtr", "GetProbability", "CollapseQubit", "Normalize", "AmplitudeWiseScalarMultiplication",
    "ExpectationValue", "ComputeNorm", "ComputeOverlap", "Print",
    # noise (pybind11/intelqs_py.cpp:260-341)
    "GetOverallSignOfChannels", "ApplyChannel", "SetNoiseTimescales", "ApplyNoiseGate",
]


def test_python_module_exposes_reference_names():
    import intelqs_py as m
    for n in ("QubitRegister", "RandomNumberGenerator", "MPIEnvironment", "EnvInit", "EnvFinalize",
              "InitializeVectorAsMaxCutCostFunction", "InitializeVectorAsWeightedMaxCutCostFunction",
              "ImplementQaoaLayerBasedOnCostFunction", "GetExpectationValueFromCostFunction",
              "CM4x4", "CM16x16", "NoisyQureg", "GetHistogramFromCostFunction",
              "GetHistogramFromCostFunctionWithWeightsRounded", "GetHistogramFromCostFunctionWithWeightsBinned"):
        assert hasattr(m, n), n
    missing = [n for n in REFERENCE_METHODS if not hasattr(m.QubitRegister, n)]
    assert not missing, missing
    assert m.MPIEnvironment.GetStateSize() == 1 and m.MPIEnvironment.GetRank() == 0


def test_rng_matches_std_mt19937_streams():
    """SetSeedStreamPtrs seeds the pool stream with `seed`: the first uniform numbers are
    those of std::mt19937(seed) through uniform_real_distribution<double>(0, 1)."""
    import intelqs_py as m
    r = m.RandomNumberGenerator()
    r.SetSeedStreamPtrs(777)
    v = r.GetUniformRandomNumbers(4, 0.0, 1.0, "pool")
    assert all(0.0 <= x < 1.0 for x in v) and len(set(v)) == 4
    r2 = m.RandomNumberGenerator()
    r2.SetSeedStreamPtrs(777)
    assert r2.GetUniformRandomNumbers(4, 0.0, 1.0, "pool") == v
