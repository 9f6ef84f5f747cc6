"""QASM front end (intel-qs_amd/bin/iqs_qasm) against the reference's own QASM interface
program (interface/src/*.cpp, built from its sources into oracle/_ref and run by
tests/golden/make_golden.py): for each script, standard output (echoed operations,
allocation messages, MeasZ probabilities) and standard error (failed and unknown
operations) must be identical."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
BIN = os.path.join(ROOT, "intel-qs_amd", "bin", "iqs_qasm")


@pytest.mark.parametrize("name", ["qasm_mixed", "qasm_ghz"])
def test_qasm_program_matches_reference(name):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} missing: build with `make -C intel-qs_amd all`")
    script = open(os.path.join(GOLDEN, name + ".qasm")).read()
    r = subprocess.run([BIN], input=script, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout == open(os.path.join(GOLDEN, name + ".out")).read()
    assert r.stderr == open(os.path.join(GOLDEN, name + ".err")).read()
