"""The oracle reproduces the committed golden fixtures (CPU).

The fixtures (tests/golden/*.npz, written by tests/golden/gen_golden.py)
freeze the pinned oracle's answers on inputs the water boxes do not reach:
C/N chemistry (TATB), the over-coordination branch, compressed bonds.  A
change to the oracle that moves any of them fails here; the GPU tests check
the HIP path against the same files.
"""
import numpy as np
import pytest

from .golden import gen_golden as G

INT_KEYS = ("far_count", "bond_count", "bond_nbr", "hbond_count", "hbond_nbr", "H_numnbrs", "iters")


@pytest.mark.parametrize("name", sorted(G.CASES))
def test_oracle_reproduces_fixture(oracle_lib, name):
    want = G.load(name)
    case = {k: want[k] for k in ("elements", "type", "x")}
    got = G.run_oracle(case)
    for k in INT_KEYS:
        assert np.array_equal(got[k], want[k]), k
    for k, v in got.items():
        if k in INT_KEYS:
            continue
        w = want[k]
        assert v.shape == w.shape, k
        scale = np.maximum(np.abs(w), 1e-12)
        assert np.all(np.abs(v - w) <= 1e-12 * scale + 1e-15), k


def test_fixtures_reach_the_branches_they_are_for():
    t = G.load("tatb_cluster")
    assert set(np.unique(t["type"]).tolist()) == {1, 2, 3, 4}           # C, H, O, N
    c = G.load("compressed")
    # over-coordinated O in O2 at 1.05 A and C in the compressed chain
    assert c["Deltap"][0] > 0 and c["Deltap"][3] > 0
    # charges neutral and bounded
    for name in G.CASES:
        z = G.load(name)
        assert abs(z["q"].sum()) < 1e-9
        assert np.all(np.abs(z["q"]) < 2.0)
