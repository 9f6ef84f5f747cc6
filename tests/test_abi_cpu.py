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


def _mt_uniform(seed, k):
    """std::mt19937(seed) through libstdc++'s generate_canonical<double, 53>: two 32-bit
    outputs per number, (g0 + g1 2^32) / 2^64 -- restated with numpy's MT19937 under the
    same (legacy) initialization."""
    import numpy as np
    bg = np.random.MT19937(0)
    bg._legacy_seeding(seed)
    raw = bg.random_raw(2 * k)
    return [(int(raw[2 * i]) + int(raw[2 * i + 1]) * 2.0 ** 32) / 2.0 ** 64 for i in range(k)]


def test_rng_streams_and_integers_match_restatement():
    """Streams seeded as src/rng_utils.cpp:84-99 (pool: seed, state: seed + 1 + state id,
    local: seed + 1 + num_states + pool rank) and integers floor(a + u (b - a))
    (src/rng_utils.cpp:170-188), bit for bit."""
    import numpy as np
    import intelqs_py as m
    r = m.RandomNumberGenerator()
    r.SetSeedStreamPtrs(777)
    assert r.GetUniformRandomNumbers(5, 0.0, 1.0, "pool") == _mt_uniform(777, 5)
    assert r.GetUniformRandomNumbers(3, -2.0, 6.0, "local") == [-2.0 + 8.0 * u for u in _mt_uniform(779, 3)]
    out = np.zeros(8, dtype=np.int32)
    r.RandomIntegersInRange(out, 8, 3, 13, "state")
    assert list(out) == [int(np.floor(3 + u * 10)) for u in _mt_uniform(778, 8)]
    buf = np.zeros(4)
    r2 = m.RandomNumberGenerator()
    r2.SetSeedStreamPtrs(777)
    r2.UniformRandomNumbers(buf, 4, 0.0, 1.0, "pool")
    assert list(buf) == _mt_uniform(777, 4)


def test_rng_counters_skip_and_shuffle():
    """Stream counters as the reference's unit test checks them
    (unit_test/include/random_number_generator_test.hpp:48-62, 154-161): n uniforms count n
    on their stream only, a Gaussian counts twice, SkipAhead(k) draws and counts k uniforms
    (src/rng_utils.cpp:101-112); ShuffleFisherYates (src/rng_utils.cpp:403-423) restated with
    integers floor(u (ra + 1)) drawn from the local stream std::mt19937(seed + 1 + num_states +
    pool rank) = std::mt19937(seed + 2), ra = n-1 .. 1."""
    import numpy as np
    import intelqs_py as m
    r = m.RandomNumberGenerator()
    r.SetSeedStreamPtrs(4242)
    r.GetUniformRandomNumbers(37, 0.0, 1.0, "local")
    assert (r.GetNumGeneratedOrSkippedLocalNumbers(), r.GetNumGeneratedOrSkippedStateNumbers(),
            r.GetNumGeneratedOrSkippedPoolNumbers()) == (37, 0, 0)
    r.GaussianRandomNumbers(np.zeros(5), 5, "pool")
    r.SkipeAhead(11, "state")
    assert (r.GetNumGeneratedOrSkippedLocalNumbers(), r.GetNumGeneratedOrSkippedStateNumbers(),
            r.GetNumGeneratedOrSkippedPoolNumbers()) == (37, 11, 10)
    # skipping k numbers = drawing k numbers
    a, b = m.RandomNumberGenerator(), m.RandomNumberGenerator()
    a.SetSeedStreamPtrs(9)
    b.SetSeedStreamPtrs(9)
    a.SkipeAhead(7, "local")
    assert a.GetUniformRandomNumbers(3, 0.0, 1.0, "local") == b.GetUniformRandomNumbers(10, 0.0, 1.0, "local")[7:]
    r.SetSeedStreamPtrs(4242)
    assert r.GetNumGeneratedOrSkippedLocalNumbers() == 0
    # shuffle: same draws, same swaps
    n = 12
    s = m.RandomNumberGenerator()
    s.SetSeedStreamPtrs(31)
    got = m.ShuffleFisherYates(list(range(n)), s, "local")
    u = _mt_uniform(33, n - 1)
    want = list(range(n))
    for i, ra in enumerate(range(n - 1, 0, -1)):
        rb = int(np.floor(0 + u[i] * (ra + 1)))
        want[ra], want[rb] = want[rb], want[ra]
    assert got == want and sorted(got) == list(range(n))
    assert s.GetNumGeneratedOrSkippedLocalNumbers() == n - 1
