"""The CPU oracle replays every golden fixture (written by the compiled reference) and must
reproduce it bit for bit: final states and read-out lines (reductions are fixed to the
reference's single-thread order)."""
import glob
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = sorted(os.path.basename(p)[:-5] for p in glob.glob(os.path.join(GOLDEN, "*.circ")))


def load(name):
    text = open(os.path.join(GOLDEN, name + ".circ")).read()
    res = [l for l in open(os.path.join(GOLDEN, name + ".res")).read().splitlines() if l]
    return text, res, np.load(os.path.join(GOLDEN, name + ".npz"))


def test_fixture_set_present():
    assert len(CASES) >= 20
    assert "brick20_dp" in CASES


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_reference_fixture(name):
    text, ref_res, g = load(name)
    st, res = oracle.run(text)
    if "state" in g.files:
        assert st.dtype == g["state"].dtype
        assert np.array_equal(st.view(np.uint8), g["state"].view(np.uint8))
    else:
        assert np.array_equal(st[g["idx"]].view(np.uint8), g["amp"].view(np.uint8))
        a = st.astype(np.complex128)
        assert abs(np.vdot(a, a).real - g["norm2"][0]) < 1e-12
        assert abs(a.sum() - g["sum"][0]) < 1e-9
    assert res == ref_res
