"""Batched evaluation: every system gives the same results alone and in a mix.

A batch of clusters and small periodic water boxes runs as one device system
(one launch per stage, per-system QEq scalars).  Each system is also run
alone.  In the batch a system is translated into its own region and binned
on another grid, so its coordinates differ in the last bit and its sums run
in another order.  The pipelined CG amplifies such last-bit differences in H
(by ~1e7: the oracle's charges move by ~4e-9 e when a system is merely
translated), and a solve can stop an iteration or two apart when its stopping
test sits at rounding level.  Per system the checks are therefore: CG
iteration counts within 2, charges within 1e-5 e at the reference tolerance
1e-6 (one iteration of truncation) and within 1e-8 e at tol 1e-10, identical
bond partners, bond orders within 1e-9 relative (a weak bond's BO = BO' -
1e-4 cancels, which turns last-bit coordinate changes into ~1e-10
relative).  One cluster is also checked against the FP64 oracle at 1e-10.
"""
import numpy as np
import pytest

from sw_reaxff_amd import batch as B
from sw_reaxff_amd import systems

from .parity import rel_err

pytestmark = pytest.mark.gpu


def mix():
    sizes = [50, 120, 333, 700, 1500]
    out = [systems.cluster(k, seed=40 + i) for i, k in enumerate(sizes)]
    out.insert(2, systems.water_box(216, seed=3))     # periodic, 648 atoms
    out.append(systems.water_box(343, seed=4))         # periodic, 1029 atoms
    return out


def bonds_of(bt, k):
    h = bt.sys
    P = bt.packed
    bi, be, bonds, BO = h.get("bond_index"), h.get("bond_end"), h.get("bonds"), h.get("BO")
    rows = range(P.row_off[k], P.row_off[k + 1])
    cnt = np.array([be[i] - bi[i] for i in rows])
    sel = np.concatenate([np.arange(bi[i], be[i]) for i in rows])
    nbr = bonds["nbr"][sel]
    # partners as (owner row - system start): ghosts map to their owner
    owner = h.get("owner")
    return cnt, owner[nbr] - P.row_off[k], BO[sel]


@pytest.mark.parametrize("tol,qtol", [(1e-10, 1e-8), (1e-6, 1e-5)])
def test_alone_equals_in_the_mix(gpu_lib, tol, qtol):
    parts = mix()
    bt = B.Batch(tol=tol)
    bt.load(parts)
    bt.evaluate(nonbonded=True, bonded=True)
    bt.sync()
    q_mix = bt.charges()
    it_mix = bt.iterations()
    e_mix = bt.energies()
    eb_mix = bt.bonded_energies()
    g_mix = [g.copy() for g in bt.gradients()]
    for k, s in enumerate(parts):
        alone = B.Batch(tol=tol)
        alone.load([s])
        alone.evaluate(nonbonded=True, bonded=True)
        alone.sync()
        q1 = alone.charges()[0]
        it1 = alone.iterations()[0]
        dq = float(np.max(np.abs(q1 - q_mix[k])))
        print(f"[batch] system {k} n={s.n} periodic={s.periodic} iters alone {tuple(it1)} mix {tuple(it_mix[k])} "
              f"max|dq| {dq:.2e}", flush=True)
        assert np.all(np.abs(np.asarray(it1) - np.asarray(it_mix[k])) <= 2), k
        assert dq <= qtol, k
        assert abs(q_mix[k].sum()) < 1e-8, k
        c1, n1, bo1 = bonds_of(alone, 0)
        c2, n2, bo2 = bonds_of(bt, k)
        assert np.array_equal(c1, c2), k
        assert np.array_equal(n1, n2), k
        assert rel_err(bo2, bo1, 1e-12) <= 1e-9, k
        # the non-bonded energy and gradient move with the charges (a charge
        # change dq moves E by ~sum_j 332 q_j / r_ij * dq ~ 1e2..1e3 dq
        # relative to E): bounds 100 x and 1000 x the charge tolerance
        e1, g1 = alone.energies()[0], alone.gradients()[0]
        de = abs(e1 - e_mix[k]) / abs(e1)
        dg = np.max(np.abs(g1 - g_mix[k])) / np.abs(g1).max()
        # bond, atom and hydrogen-bond energies do not depend on the charges (but for e_pol)
        eb1 = alone.bonded_energies()[0]
        sel = [0, 1, 2, 3, 5, 6, 7, 8, 9, 10]
        deb = np.abs(eb1[sel] - eb_mix[k][sel]) / np.maximum(np.abs(eb1[sel]), 1e-3)
        dpol = abs(eb1[4] - eb_mix[k][4]) / max(abs(eb1[4]), 1.0)
        print(f"[batch] system {k}: E_nb {e1:.10e} rel diff {de:.2e}, bond/atom {eb1} rel diff {deb.max():.2e} "
              f"pol {dpol:.2e}, gradient rel diff {dg:.2e}", flush=True)
        assert de <= 100 * qtol, k
        assert dg <= 1000 * qtol, k
        assert deb.max() <= 1e-10, k
        assert dpol <= 100 * qtol, k


def test_mixed_batch(gpu_lib):
    """Mixed-precision QEq over a batch (per-system outer and inner scalars):
    every system's charges lie within 1e-6 e of the FP64 batch solved to
    1e-10, and equal its charges alone within 1e-7 e (alone, the single-
    system kernels run; the FP32 rounding of a translated H entry can flip
    its last bit), with residual + inner iteration counts within 3."""
    parts = mix()
    ref = B.Batch(tol=1e-10)
    ref.load(parts)
    ref.evaluate()
    ref.sync()
    q_ref = [q.copy() for q in ref.charges()]
    bt = B.Batch(tol=1e-6, mixed=True)
    bt.load(parts)
    bt.evaluate()
    bt.sync()
    q_mix = bt.charges()
    it_mix = bt.iterations()
    for k, s in enumerate(parts):
        alone = B.Batch(tol=1e-6, mixed=True)
        alone.load([s])
        alone.evaluate()
        alone.sync()
        q1 = alone.charges()[0]
        it1 = alone.iterations()[0]
        da = float(np.max(np.abs(q1 - q_mix[k])))
        dr = float(np.max(np.abs(q_mix[k] - q_ref[k])))
        print(f"[batch mixed] system {k} n={s.n} periodic={s.periodic} matvecs alone {tuple(it1)} mix "
              f"{tuple(it_mix[k])} max|dq| alone-mix {da:.2e} vs FP64(1e-10) {dr:.2e}", flush=True)
        assert np.all(np.abs(np.asarray(it1) - np.asarray(it_mix[k])) <= 3), k
        assert da <= 1e-7, k
        assert dr <= 1e-6, k
        assert abs(q_mix[k].sum()) < 1e-8, k


def test_batch_cluster_against_oracle(gpu_lib, oracle_lib, ffield):
    parts = mix()
    bt = B.Batch(tol=1e-10)
    P = bt.load(parts)
    bt.evaluate()
    bt.sync()
    k = 3                                   # the 333-atom cluster
    s = parts[k]
    o = oracle_lib.Oracle(ffield, B.ELEMENTS, tol=1e-10, threads=4)
    tmap = np.array([B.ELEMENTS.index(e) + 1 for e in s.elements], dtype=np.int32)
    o.set_atoms(s.n, s.x, tmap[s.type - 1], s.tag, np.arange(s.n, dtype=np.int32))
    o.build_verlet(12.0)
    o.step()
    dq = np.max(np.abs(bt.charges()[k] - o.get("q")[:s.n]))
    assert dq <= 1e-9
    bi, be = o.get("bond_index"), o.get("bond_end")
    c2, _, bo2 = bonds_of(bt, k)
    assert np.array_equal(be - bi, c2)
    obo = np.concatenate([o.get("BO")[bi[i]:be[i]] for i in range(s.n)])
    assert rel_err(bo2, obo, 1e-12) <= 1e-10


def test_batch_rejects_bad_offsets(gpu_lib):
    import ctypes as C
    bt = B.Batch()
    bt.load(mix()[:2])
    h = bt.sys
    bad = np.array([0, 10, 5, bt.packed.n], dtype=np.int32)
    rc = h.L.swr_set_batch(h.h, 3, bad.ctypes.data_as(C.POINTER(C.c_int)))
    assert rc == 1
    short = np.array([0, 7], dtype=np.int32)
    assert h.L.swr_set_batch(h.h, 1, short.ctypes.data_as(C.POINTER(C.c_int))) == 1
