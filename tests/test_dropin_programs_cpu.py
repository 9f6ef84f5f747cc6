"""Source compatibility of the drop-in for the reference's own C++ callers.

Every program under the reference's examples/, tutorials/ and benchmarks/ is compiled,
unchanged and from where it lies, against the drop-in headers (intel-qs_amd/include) and
linked against lib/libiqs_amd.so (oracle/Makefile, target `progs`).  No GPU is needed to
compile and link; running them is the GPU test tests/test_gpu_reference_programs.py.
Skipped where the reference sources are absent.
"""
import glob
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
DIRS = ("examples", "tutorials", "benchmarks")

needs_ref = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "examples")),
                               reason="reference sources absent")


def _programs():
    return sorted(os.path.splitext(os.path.basename(f))[0]
                  for d in DIRS for f in glob.glob(os.path.join(REF, d, "*.cpp")))


@needs_ref
def test_every_reference_program_compiles_and_links_against_the_dropin(tmp_path):
    lib = os.path.join(ROOT, "intel-qs_amd", "lib", "libiqs_amd.so")
    if not os.path.exists(lib):
        pytest.skip("lib/libiqs_amd.so not built")
    names = _programs()
    assert len(names) >= 13, names
    targets = [str(tmp_path / "dropin" / n) for n in names]
    r = subprocess.run(["make", "-k", "-j8", "-C", os.path.join(ROOT, "oracle"), f"PROGDIR={tmp_path}"] + targets,
                       capture_output=True, text=True, timeout=600)
    missing = [n for n, t in zip(names, targets) if not os.path.exists(t)]
    assert not missing, f"failed to build {missing}:\n{r.stderr[-4000:]}"
    # linked against the drop-in, not the reference
    out = subprocess.run(["ldd", targets[0]], capture_output=True, text=True).stdout
    assert "libiqs_amd.so" in out and "libiqs_ref" not in out
