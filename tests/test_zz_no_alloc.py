"""swr_qeq_cg_noalloc, the FP64 QEq solve that allocates nothing, against
swr_qeq_cg on the same H, right-hand sides and warm start.

Both run the same kernels in the same order on the same data (only the
buffers differ: the system's own for swr_qeq_cg, the library's static
per-GPU scratch array for the other), so the solutions agree to the bit and
the iteration counts are equal; the bar asserted is the charges' 1e-6 e
tolerance and 1e-12 on s and t.  Two calls in a row share the scratch, the
second on a smaller system, so stale contents of the first would show.
Device buffers for b and x come from the HIP runtime the library itself
links (the test allocates them; the entry point does not).
"""
import ctypes as C

import numpy as np
import pytest

from sw_reaxff_amd import systems

from .parity import harness

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    hip = C.CDLL("libamdhip64.so.7")   # the copy libswreax.so already loaded
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


def _ok(rc, what):
    assert rc == 0, f"{what}: hipError {rc}"


def _solve_both(h, hip):
    """(x from swr_qeq_cg, its iterations, x from swr_qeq_cg_noalloc, its
    iterations), x as (n, 2) arrays of (s, t)."""
    from sw_reaxff_amd import _lib
    h.call("qeq_init_matvec")                      # right-hand sides and the warm start
    h.sync()
    b = np.ascontiguousarray(np.stack([h.get("b_s"), h.get("b_t")], axis=1))
    x0 = np.ascontiguousarray(np.stack([h.get("s"), h.get("t")], axis=1))
    its_ref = h.qeq_cg()
    x_ref = np.stack([h.get("s"), h.get("t")], axis=1)
    db, dx = C.c_void_p(), C.c_void_p()
    _ok(hip.hipMalloc(C.byref(db), b.nbytes), "hipMalloc b")
    _ok(hip.hipMalloc(C.byref(dx), x0.nbytes), "hipMalloc x")
    try:
        _ok(hip.hipMemcpy(db, b.ctypes.data, b.nbytes, H2D), "copy b")
        _ok(hip.hipMemcpy(dx, x0.ctypes.data, x0.nbytes, H2D), "copy x")
        a, c = C.c_int(0), C.c_int(0)
        _lib.check(h.L.swr_qeq_cg_noalloc(h.h, db, dx, C.byref(a), C.byref(c)), "swr_qeq_cg_noalloc")
        x = np.empty_like(x0)
        _ok(hip.hipMemcpy(x.ctypes.data, dx, x.nbytes, D2H), "copy x back")
    finally:
        hip.hipFree(db)
        hip.hipFree(dx)
    return x_ref, its_ref, x, (a.value, c.value)


def _charges(x):
    s, t = x[:, 0], x[:, 1]
    return s - s.sum() / t.sum() * t


def test_noalloc_cg_matches_qeq_cg_twice_second_smaller(gpu_lib, ffield):
    hip = _hip()
    first = harness(systems.water_box(648, seed=41), ffield, tol=1e-10)
    second = harness(systems.water_box(216, seed=42), ffield, tol=1e-10)
    for name, lmp in (("1944 atoms", first), ("648 atoms", second)):
        h = lmp.system
        x_ref, its_ref, x, its = _solve_both(h, hip)
        dx = float(np.abs(x - x_ref).max())
        dq = float(np.abs(_charges(x) - _charges(x_ref)).max())
        print(f"[no-alloc] {name}: iterations {its} (swr_qeq_cg {its_ref}) max|dx| {dx:.2e} max|dq| {dq:.2e}",
              flush=True)
        assert its == its_ref
        assert dx <= 1e-12 and dq <= 1e-6


def test_noalloc_cg_rejects_what_it_cannot_hold(gpu_lib, ffield):
    from sw_reaxff_amd import _lib
    hip = _hip()
    # beyond the scratch's 32768 rows: refused before any launch
    big = harness(systems.water_box(11000, seed=43), ffield)
    assert big.system.scalar("n") > 32768
    with pytest.raises(_lib.SwrError) as ei:
        _solve_both(big.system, hip)
    assert ei.value.code == 1 and "at most 32768" in str(ei.value)   # SWR_ERR_ARG
    # the mixed-precision solve is not this entry point's
    mixed = harness(systems.water_box(216, seed=44), ffield, mixed=True)
    with pytest.raises(_lib.SwrError) as ei:
        _solve_both(mixed.system, hip)
    assert ei.value.code == 1 and "FP64" in str(ei.value)
