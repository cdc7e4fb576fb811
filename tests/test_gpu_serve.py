"""Serving loop: every served result equals the same evaluation done alone.

A short stream of requests (new clusters and periodic water boxes, some kept
resident and stepped by follow-on requests) runs through the server with two
overlapping workers.  Afterwards each new system is evaluated alone and each
resident system is replayed alone step by step.

Tolerances: a batched system is translated into its own region and binned on
another grid, so its sums run in another order; the pipelined CG amplifies
those last-bit differences (~1e7) and can stop an iteration or two apart
from the lone run, which moves charges by up to the solver's own truncation
(~5e-6 e at the reference tolerance 1e-6).  Served charges must therefore
agree with the lone run within 1e-5 e at tol 1e-6 and within 1e-8 e at tol
1e-10; with mixed-precision QEq (per-system FP64 refinement to tol/1000
around the FP32 inner solve) within 1e-7 e.

Batched residents: each step is one velocity-Verlet step with the ReaxFF
forces, so its positions must be exactly the last state's advanced with the
forces served for it (checked bit for bit: host arithmetic), evaluated in a
batch with the warm start the last step left in device memory; the replay
evaluates the same positions alone, chaining its own warm starts the same
way: charges to the same tolerance, forces within 1000x the charge
tolerance of the largest force component (the charges enter them).  Device
residents (their own device system with fix nve and forces, set up at the
first follow-on step from the s/t histories the batched solve left) are
replayed from the same histories and stepped the same way, so their step
charges agree to 1e-12 e.
"""
import numpy as np
import pytest

from sw_reaxff_amd import serve
from sw_reaxff_amd.batch import Batch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tol,qtol,mixed,mode", [(1e-6, 1e-5, False, "batched"), (1e-10, 1e-8, False, "batched"),
                                                  (1e-6, 1e-7, True, "batched"), (1e-6, 1e-5, False, "device")])
def test_served_results_equal_alone(gpu_lib, ffield, tol, qtol, mixed, mode):
    # arrivals faster than one evaluation, so batches of several systems form
    reqs = serve.make_requests(30, rate=5000.0, seed=11, lo=20, hi=1500, frac_step=0.3, frac_resident=0.25)
    _check_served(reqs, ffield, tol, qtol, mixed, mode, lanes=2)


@pytest.mark.timeout(900)
def test_c4_serving_stream(gpu_lib, ffield):
    """The serving workload's own mix: sizes log-uniform on [20, 20000], half
    periodic water boxes and half clusters, 30% follow-on MD steps, Poisson
    arrivals, mixed-precision QEq at tol 1e-6, 200 GB admission, four lanes."""
    reqs = serve.make_requests(80, rate=1500.0, seed=13, lo=20, hi=20000, frac_step=0.3, frac_resident=0.2)
    assert max(r.system.n for r in reqs if r.kind == "new") > 5000
    _check_served(reqs, ffield, 1e-6, 1e-7, True, "batched", lanes=4)


def _check_served(reqs, ffield, tol, qtol, mixed, mode, lanes):
    srv = serve.Server(lanes=lanes, hbm_budget=200e9, ffield=ffield, tol=tol, mixed=mixed, resident=mode)
    srv.run(reqs)
    assert all(r.q is not None for r in reqs)
    assert any(r.kind == "step" for r in reqs) and any(r.batch_size > 1 for r in reqs)
    print(f"[serve] {mode} {serve.summary(reqs, max(r.done for r in reqs))}", flush=True)
    alone = Batch(ffield=ffield, tol=tol, mixed=mixed)
    worst = 0.0
    for r in reqs:
        if r.kind == "new":
            alone.load([r.system])
            alone.evaluate()
            alone.sync()
            dq = float(np.max(np.abs(alone.charges()[0] - r.q)))
            worst = max(worst, dq)
            its = tuple(int(v) for v in alone.iterations()[0])
            if dq > 1e-7:
                print(f"[serve] request {r.rid} n={r.system.n} batch {r.batch_size}: max|dq| {dq:.2e}, "
                      f"iterations served {r.iters} alone {its}", flush=True)
            assert dq <= qtol, (r.rid, r.system.n, r.batch_size, dq, r.iters, its)
    print(f"[serve] tol {tol}: worst served-vs-alone charge difference {worst:.2e}", flush=True)
    # residents: replay alone in the order the server stepped them
    worst = 0.0
    for rid in sorted({r.resident for r in reqs if r.resident >= 0}):
        first = next(r for r in reqs if r.kind == "new" and r.resident == rid)
        steps = sorted((r for r in reqs if r.kind == "step" and r.resident == rid), key=lambda r: r.step_index)
        if mode == "device":
            assert first.hist is not None
            res = serve._Resident(first.system, ffield, tol, mixed, hist=first.hist)
            for k, r in enumerate(steps, 1):
                assert r.step_index == k
                assert np.max(np.abs(res.step() - r.q)) <= 1e-12, (rid, k)
            continue
        alone.load([first.system])
        alone.evaluate()
        store = alone.save_history(0, fill=True)
        alone.sync()
        prev, prev_f = first.system, first.f
        assert prev_f is not None and prev_f.shape == (first.system.n, 3)
        for k, r in enumerate(steps, 1):
            assert r.step_index == k
            want = serve.nve_initial(prev, prev_f, srv.dt)
            assert np.array_equal(r.system.x, want.x) and np.array_equal(r.system.v, want.v), (rid, k)
            alone.load([r.system], hist=[store])
            alone.evaluate(nonbonded=True, bonded=True)
            alone.save_history(0, store)
            alone.sync()
            dq = float(np.max(np.abs(alone.charges()[0] - r.q)))
            f1 = -alone.gradients()[0]
            df = float(np.max(np.abs(f1 - r.f)) / np.abs(f1).max())
            worst = max(worst, dq)
            assert dq <= qtol, (rid, k, dq)
            assert df <= 1000 * qtol, (rid, k, df)
            prev, prev_f = serve.nve_final(r.system, r.f, srv.dt), r.f
    print(f"[serve] {mode} tol {tol}: worst resident step served-vs-alone charge difference {worst:.2e}", flush=True)
