"""The autotuner of launch shapes: it picks one of the compiled shapes of
each family, and every shape computes the same results -- the far pass the
same lists bit for bit, the SpMV shapes the same charges up to the order of
the sums inside an H row (1e-9 e at tol 1e-10 for the FP64 solve, 1e-7 e for
the mixed-precision one, whose FP32 inner sums round differently)."""
import numpy as np
import pytest

from .parity import harness, water_c0

pytestmark = pytest.mark.gpu

SPMV = ["T512G8U4", "T512G8U2", "T256G8U4", "T256G4U4"]
FAR = ["G4W2", "G4W4", "G6W2", "G8W4"]


def _resolve(h, hist):
    """QEq again from the same warm start (init_matvec, CG, calculate_Q)."""
    h.put("s_hist", hist[0])
    h.put("t_hist", hist[1])
    for st in ("qeq_init_matvec",):
        h.call(st)
    its = h.qeq_cg()
    h.call("calculate_q")
    h.sync()
    return h.get("q")[: h.scalar("n")].copy(), its


def test_autotune_picks_a_shape_per_family(gpu_lib, ffield):
    lmp = harness(water_c0(seed=31), ffield)
    h = lmp.system
    rep = h.autotune(reps=3)
    print(f"[tune] {rep}", flush=True)
    assert rep["atoms"] == 1944
    assert rep["spmv_fp64"]["chosen"] in SPMV and rep["far"]["chosen"] in FAR
    assert set(rep["spmv_fp64"]["us"]) == set(SPMV) and set(rep["far"]["us"]) == set(FAR)
    assert all(v > 0 for v in rep["spmv_fp64"]["us"].values())
    assert h.shape("spmv_fp64") == rep["spmv_fp64"]["chosen"] and h.shape("far") == rep["far"]["chosen"]
    lmp.run(3)   # the run continues with the chosen shapes


def test_far_shapes_give_identical_lists(gpu_lib, ffield):
    lmp = harness(water_c0(seed=32), ffield)
    h = lmp.system
    keys = ("far", "far_index", "far_end", "bonds", "BO", "hbonds", "hbond_end", "bond_mark", "num_bonds")
    ref = None
    for shape in FAR:
        h.set_shape("far", shape)
        h.call("pair_compute")
        h.sync()
        got = {k: h.get(k) for k in keys}
        if ref is None:
            ref = got
            continue
        for k in keys:
            if k == "hbonds":   # ptr holds the device address of the far entry: same list, same addresses
                assert np.array_equal(got[k]["nbr"], ref[k]["nbr"]) and np.array_equal(got[k]["ptr"], ref[k]["ptr"])
            else:
                assert np.array_equal(got[k], ref[k]), (shape, k)


@pytest.mark.parametrize("mixed", [False, True])
def test_spmv_shapes_give_the_same_charges(gpu_lib, ffield, mixed):
    lmp = harness(water_c0(seed=33), ffield, tol=1e-10 if not mixed else 1e-6, mixed=mixed)
    h = lmp.system
    hist = (h.get("s_hist").copy(), h.get("t_hist").copy())
    fams = ["spmv_fp64"] if not mixed else ["spmv_fp32", "spmv_resid"]
    q0, its0 = _resolve(h, hist)
    atol = 1e-9 if not mixed else 1e-7
    for fam in fams:
        for shape in SPMV:
            h.set_shape(fam, shape)
            q, its = _resolve(h, hist)
            dq = float(np.max(np.abs(q - q0)))
            print(f"[tune] {fam} {shape}: iterations {its} (default {its0}) max|dq| {dq:.2e}", flush=True)
            assert dq <= atol, (fam, shape, dq)
        h.set_shape(fam, "T512G8U4")
    if mixed:
        rep = h.autotune(reps=3)
        assert rep["spmv_fp32"]["chosen"] in SPMV and rep["spmv_resid"]["chosen"] in SPMV
