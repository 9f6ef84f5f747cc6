"""The reference's own example and tutorial programs, unchanged, run through the drop-in on
the GPU and through the compiled reference on the host: their printed output must agree.

Both builds come from oracle/Makefile (target `progs`, made by build() where the reference
sources exist): oracle/_ref/progs/ref/<name> links the reference compiled from its sources,
oracle/_ref/progs/dropin/<name> the same source against intel-qs_amd/include and
lib/libiqs_amd.so.  Output is compared line by line: the text between numbers must be
identical up to the width of the runs of blanks (a printed -0.00000000 is one column wider
than 0.00000000), and numbers must agree to 1e-6 (absolute, plus 1e-6 relative).  Numbers on lines
that report an error against another method (the QFT-vs-FFT lines) only have to be below
1e-5 in both runs, since they measure round-off.  Timing and thread-count lines are
dropped.  References: examples/*.cpp, tutorials/*.cpp, benchmarks/specv2_bench.cpp.
"""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROGS = os.path.join(ROOT, "oracle", "_ref", "progs")
NUM = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?")
DROP = re.compile(r"time|seconds|OMP number of threads|Without MKL", re.IGNORECASE)

CASES = [
    ("quantum_fourier_transform", ["12"]),
    ("grover_4qubit", []),
    ("expect_value_test", []),
    ("test_of_custom_gates", ["6"]),
    ("get_started_with_IQS", []),
    ("heisenberg_dynamics", ["10"]),
    ("communication_reduction_via_qubit_reordering", ["22"]),
    ("circuit_with_noise_gates", ["6"]),
    ("get_started_with_noisy_IQS", []),
    ("benchgates", ["10"]),
    ("specv2_bench", ["12", "1"]),
]


def _run(path, args):
    env = dict(os.environ, OMP_NUM_THREADS="4")
    r = subprocess.run([path] + args, capture_output=True, text=True, timeout=240, env=env)
    lines = [l.rstrip() for l in r.stdout.splitlines() if not DROP.search(l)]
    return r.returncode, lines, r.stderr


def outputs_agree(got, want):
    """None if the two outputs agree (see module docstring), else the first difference."""
    if len(got) != len(want):
        return f"{len(got)} lines vs {len(want)}"
    for k, (g, w) in enumerate(zip(got, want)):
        if NUM.sub("#", g).split() != NUM.sub("#", w).split():
            return f"line {k}: {g!r} vs {w!r}"
        err_line = re.search(r"error|diff", w, re.IGNORECASE)
        for a, b in zip(NUM.findall(g), NUM.findall(w)):
            a, b = float(a), float(b)
            if abs(a - b) <= 1e-6 + 1e-6 * abs(b):
                continue
            if err_line and abs(a) < 1e-5 and abs(b) < 1e-5:
                continue
            return f"line {k}: {g!r} vs {w!r}"
    return None


def test_comparator_catches_differences():
    a = ["<Z> on qubit 0: 0.866025624917", "x (1,0)"]
    assert outputs_agree(a, a) is None
    assert outputs_agree(["<Z> on qubit 0: 0.866035"], ["<Z> on qubit 0: 0.866025624917"]) is not None
    assert outputs_agree(["<Z> on qubit 1: 0.866025624917"], ["<Z> on qubit 0: 0.866025624917"]) is not None
    assert outputs_agree(["psi = <1|psi> = (1,0)"], ["psi = <1|psi> = (0,1)"]) is not None
    assert outputs_agree(["error max(absdiff: 3e-08)"], ["error max(absdiff: 7e-08)"]) is None
    assert outputs_agree(["error max(absdiff: 3e-03)"], ["error max(absdiff: 7e-08)"]) is not None
    assert outputs_agree(a, a[:1]) is not None
    assert outputs_agree(["\t0.7    + i * -0.00000000     % |1>"], ["\t0.7    + i * 0.00000000      % |1>"]) is None
    assert outputs_agree(["\t0.7    + i * 0.0     % |1>"], ["\t0.7    + i * 0.0     % |0>"]) is not None


@pytest.mark.gpu
@pytest.mark.parametrize("name,args", CASES, ids=[c[0] for c in CASES])
def test_reference_program_output(name, args):
    ref, dropin = os.path.join(PROGS, "ref", name), os.path.join(PROGS, "dropin", name)
    if not (os.path.exists(ref) and os.path.exists(dropin)):
        pytest.skip("reference programs not built (build() builds them where the reference sources exist)")
    print(f"reference program {name} {' '.join(args)}", flush=True)
    rc_w, want, err_w = _run(ref, args)
    rc_g, got, err_g = _run(dropin, args)
    assert rc_g == rc_w, (rc_g, rc_w, err_g[-2000:])
    assert len(want) > 0
    diff = outputs_agree(got, want)
    assert diff is None, diff
