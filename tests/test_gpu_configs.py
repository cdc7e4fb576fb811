"""GPU parity on the benchmark's own workloads, against the FP64 oracle.

  c0  the 648-water box, 100 NVE steps with forces: every step's lists,
      bond orders, charges, energies and forces.  At the reference tolerance
      (1e-6) the oracle is fed each step's GPU state including the QEq history
      the step started from; at 1e-10 the oracle keeps its own history (an
      independent warm-start chain over the same positions).
  c1  the 32,000-water box (96,000 atoms): setup and 3 MD steps, whole
      state against the oracle at tol 1e-6, and setup at tol 1e-10.
  c2  mixed-precision QEq: 96,000 atoms against the converged FP64 oracle
      for 3 steps (charges within 1e-6 e); 768,000 atoms against the GPU's
      FP64 path at tol 1e-10 for 5 steps.
  c3  the 512 water/methanol clusters (50-3000 atoms) in one batch:
      alone == in the batch for 32 sampled clusters, and 3 clusters against
      the oracle.

Tolerances: lists equal (bitwise where both sides evaluate one expression
in one order), bond orders and coefficients 1e-10 relative, charges 1e-6 e
at tol 1e-6 and 1e-9 e at tol 1e-10, energies and forces 1e-8 relative at
tol 1e-10 (the charges enter them).
"""
import numpy as np
import pytest

from sw_reaxff_amd import batch as B
from sw_reaxff_amd import systems

from .parity import (Q_ATOL, compare_energy_forces, compare_state, feed_oracle, harness, history_of, rel_err,
                     atoms_of)

pytestmark = pytest.mark.gpu

THREADS = 16


def _oracle(oracle_lib, ffield, elements, tol):
    return oracle_lib.Oracle(ffield, elements, tol=tol, threads=THREADS)


# ------------------------------------------------------------------ c0
@pytest.mark.timeout(900)
def test_c0_100_steps_reference_tolerance(gpu_lib, oracle_lib, ffield):
    s = systems.water_box(648, seed=1)
    lmp = harness(s, ffield, forces=True)
    h = lmp.system
    o = _oracle(oracle_lib, ffield, s.elements, 1e-6)
    dq, dit, dbo, de, df = [], [], [], [], []
    hist = None
    for step in range(101):
        if step:
            hist = history_of(h)
            lmp.run(1)
        A = feed_oracle(o, h, hist)
        its = o.step()
        st = compare_state(h, o, Q_ATOL)
        git = (h.scalar("matvecs_s"), h.scalar("matvecs_t"))
        e, f = compare_energy_forces(h, o, A["owner"], A["n"])
        dq.append(st["dq"])
        dit.append(max(abs(git[0] - its[0]), abs(git[1] - its[1])))
        dbo.append(st["bo_rel"])
        de.append(e)
        df.append(f)
        assert dit[-1] <= 1, (step, git, its)
    print(f"[c0 tol 1e-6] 101 states: max|dq| max {max(dq):.2e} median {np.median(dq):.2e}; "
          f"iteration differences {np.bincount(dit).tolist()}; BO rel {max(dbo):.2e}; "
          f"energy rel {max(de):.2e}; force rel max {max(df):.2e} median {np.median(df):.2e}", flush=True)
    assert max(df) <= 1e-5


@pytest.mark.timeout(900)
def test_c0_100_steps_tight_tolerance(gpu_lib, oracle_lib, ffield):
    s = systems.water_box(648, seed=1)
    lmp = harness(s, ffield, tol=1e-10, forces=True)
    h = lmp.system
    o = _oracle(oracle_lib, ffield, s.elements, 1e-10)
    dq, de, df = [], [], []
    for step in range(101):
        if step:
            lmp.run(1)
        A = feed_oracle(o, h)          # the oracle's own history chain
        o.step()
        st = compare_state(h, o, 1e-9, far=(step % 10 == 0))
        e, f = compare_energy_forces(h, o, A["owner"], A["n"])
        dq.append(st["dq"])
        de.append(e)
        df.append(f)
    print(f"[c0 tol 1e-10] 101 states: max|dq| {max(dq):.2e}; energy rel {max(de):.2e}; "
          f"force rel {max(df):.2e}", flush=True)
    assert max(de) <= 1e-8
    assert max(df) <= 1e-8


# ------------------------------------------------------------------ c1
@pytest.mark.timeout(1200)
def test_c1_water96k_setup_and_steps(gpu_lib, oracle_lib, ffield):
    s = systems.water_box(32000, seed=2024)
    lmp = harness(s, ffield, forces=True)
    h = lmp.system
    print("[c1] GPU setup done", flush=True)
    o = _oracle(oracle_lib, ffield, s.elements, 1e-6)
    # the Verlet list as sets against the oracle's own cell list
    A = atoms_of(h)
    n, N = A["n"], A["N"]
    o.set_atoms(n, A["x"], A["type"], A["tag"], A["owner"])
    o.build_verlet(12.0)
    off, lst = h.get("verlet_offsets"), h.get("verlet_list")
    ooff, olst = o.get("verlet_offsets"), o.get("verlet_list")
    assert np.array_equal(np.diff(off), np.diff(ooff))
    gi = np.repeat(np.arange(N, dtype=np.int64), np.diff(off))
    key_g = np.sort(gi * N + lst[:off[-1]])
    key_o = np.sort(gi * N + olst[:ooff[-1]])
    assert np.array_equal(key_g, key_o)
    del key_g, key_o, gi, lst, olst
    print("[c1] Verlet list equals the oracle's as sets", flush=True)
    hist = None
    for step in range(4):
        if step:
            hist = history_of(h)
            lmp.run(1)
        A = feed_oracle(o, h, hist)
        its = o.step()
        st = compare_state(h, o, Q_ATOL)
        git = (h.scalar("matvecs_s"), h.scalar("matvecs_t"))
        e, f = compare_energy_forces(h, o, A["owner"], A["n"])
        print(f"[c1] state {step}: n {n} N {A['N']} far {st['far']} bonds {st['bonds']} hbonds {st['hbonds']} "
              f"max|dq| {st['dq']:.2e} BO rel {st['bo_rel']:.2e} iterations gpu {git} oracle {its} "
              f"energy rel {e:.2e} force rel {f:.2e}", flush=True)
        assert abs(git[0] - its[0]) <= 1 and abs(git[1] - its[1]) <= 1
        assert f <= 1e-5


@pytest.mark.timeout(600)
def test_c1_water96k_tight_tolerance(gpu_lib, oracle_lib, ffield):
    s = systems.water_box(32000, seed=2024)
    lmp = harness(s, ffield, tol=1e-10)
    h = lmp.system
    o = _oracle(oracle_lib, ffield, s.elements, 1e-10)
    feed_oracle(o, h)
    o.step()
    st = compare_state(h, o, 1e-9, far=False)
    print(f"[c1 tol 1e-10] setup max|dq| {st['dq']:.2e}", flush=True)


# ------------------------------------------------------------------ c2
@pytest.mark.timeout(1200)
def test_c2_mixed_water96k_against_converged_oracle(gpu_lib, oracle_lib, ffield):
    s = systems.water_box(32000, seed=2024)
    lmp = harness(s, ffield, mixed=True, forces=True)
    h = lmp.system
    o = _oracle(oracle_lib, ffield, s.elements, 1e-10)
    for step in range(4):
        if step:
            lmp.run(1)
        A = feed_oracle(o, h, history_of(h))
        o.step()
        st = compare_state(h, o, 1e-6, far=False)
        e, f = compare_energy_forces(h, o, A["owner"], A["n"])
        print(f"[c2 96k mixed] state {step}: max|dq| vs converged FP64 {st['dq']:.2e} energy rel {e:.2e} "
              f"force rel {f:.2e} matvecs {h.scalar('matvecs_s')},{h.scalar('matvecs_t')}", flush=True)
        assert f <= 1e-5


@pytest.mark.timeout(900)
def test_c2_mixed_water768k_against_fp64(gpu_lib, ffield):
    s = systems.water_box(256000, seed=2024)
    mixed = harness(s, ffield, mixed=True)
    fp64 = harness(s, ffield, tol=1e-10)
    hm, hf = mixed.system, fp64.system
    n = s.n
    worst = 0.0
    for step in range(6):
        if step:
            mixed.run(1)
            fp64.run(1)
        assert np.array_equal(hm.get("x")[:3 * n], hf.get("x")[:3 * n])
        qm, qf = hm.get("q")[:n], hf.get("q")[:n]
        dq = float(np.max(np.abs(qm - qf)))
        worst = max(worst, dq)
        print(f"[c2 768k mixed] state {step}: max|dq| vs FP64 tol 1e-10 {dq:.2e} sum q {qm.sum():.1e} "
              f"matvecs {hm.scalar('matvecs_s')},{hm.scalar('matvecs_t')}", flush=True)
        assert dq <= 1e-6
        assert abs(qm.sum()) < 1e-8
    mixed.close()
    fp64.close()


# ------------------------------------------------------------------ c3
def _clusters512():
    sizes = systems.cluster_sizes(512, 50, 3000, seed=2024)
    return [systems.cluster(int(k), seed=2025 + i) for i, k in enumerate(sizes)]


@pytest.fixture(scope="module")
def c3_batch():
    parts = _clusters512()
    print(f"[c3] built {len(parts)} clusters", flush=True)
    bt = B.Batch(tol=1e-6)
    bt.load(parts)
    bt.evaluate(nonbonded=True, bonded=True)
    bt.sync()
    return parts, bt


@pytest.mark.timeout(900)
def test_c3_alone_equals_in_the_batch(gpu_lib, c3_batch):
    parts, bt = c3_batch
    q_mix, it_mix = bt.charges(), bt.iterations()
    e_mix, eb_mix, g_mix = bt.energies(), bt.bonded_energies(), bt.gradients()
    rng = np.random.default_rng(7)
    sample = sorted(set(rng.choice(len(parts), 30, replace=False).tolist())
                    | {int(np.argmin([p.n for p in parts])), int(np.argmax([p.n for p in parts]))})
    assert len(sample) >= 32
    alone = B.Batch(tol=1e-6)
    worst = {"dq": 0.0, "de": 0.0, "dg": 0.0}
    for k in sample:
        alone.load([parts[k]])
        alone.evaluate(nonbonded=True, bonded=True)
        alone.sync()
        dq = float(np.max(np.abs(alone.charges()[0] - q_mix[k])))
        assert np.all(np.abs(alone.iterations()[0] - it_mix[k]) <= 2), k
        assert dq <= 1e-5, (k, dq)
        e1 = np.concatenate([[alone.energies()[0]], alone.bonded_energies()[0]])
        em = np.concatenate([[e_mix[k]], eb_mix[k]])
        de = float(np.max(np.abs(e1 - em) / np.maximum(np.abs(e1), 1.0)))
        g1 = alone.gradients()[0]
        dg = float(np.max(np.abs(g1 - g_mix[k])) / np.abs(g1).max())
        assert de <= 1e-3 and dg <= 1e-2, (k, de, dg)
        worst = {"dq": max(worst["dq"], dq), "de": max(worst["de"], de), "dg": max(worst["dg"], dg)}
    print(f"[c3] {len(parts)} clusters, {sum(p.n for p in parts)} atoms; {len(sample)} sampled alone vs "
          f"batched: {worst}", flush=True)


@pytest.mark.timeout(900)
def test_c3_clusters_against_oracle(gpu_lib, oracle_lib, ffield):
    """The batch at tol 1e-10 against the oracle on the smallest, a median
    and the largest cluster: bond orders 1e-10, every energy term and the
    gradient 1e-8 relative.  A cold-started solve to tol 1e-10 leaves ~1e-9 e
    of truncation that differs with the summation order (the pipelined CG
    amplifies last-bit differences), so the charges of both sides are
    checked against a dense solve of the oracle's H: each within 1e-8 e of
    it, and within 1e-8 e of each other."""
    parts = _clusters512()
    print(f"[c3 oracle] built {len(parts)} clusters", flush=True)
    bt = B.Batch(tol=1e-10)
    bt.load(parts)
    bt.evaluate(nonbonded=True, bonded=True)
    bt.sync()
    sizes = np.array([p.n for p in parts])
    picks = [int(np.argmin(sizes)), int(np.argsort(sizes)[len(sizes) // 2]), int(np.argmax(sizes))]
    h = bt.sys
    P = bt.packed
    bi, be, BO = h.get("bond_index"), h.get("bond_end"), h.get("BO")
    voff, vlist = h.get("verlet_offsets"), h.get("verlet_list")
    q, eb, en, grads = bt.charges(), bt.bonded_energies(), bt.energies(), bt.gradients()
    for k in picks:
        s = parts[k]
        r = bt.rows(k)
        # the oracle gets the cluster as the batch holds it: translated into
        # its cell and with the batch's Verlet rows (clusters have no ghosts),
        # so both sides sum in one order over bit-identical coordinates
        o = oracle_lib.Oracle(ffield, B.ELEMENTS, tol=1e-10, threads=THREADS)
        o.set_atoms(s.n, P.x[r], P.type[r], P.tag[r], np.arange(s.n, dtype=np.int32))
        assert np.all(vlist[voff[r.start]:voff[r.stop]] >= r.start) and np.all(vlist[voff[r.start]:voff[r.stop]] < r.stop)
        o.set_verlet(voff[r.start:r.stop + 1] - voff[r.start], vlist[voff[r.start]:voff[r.stop]] - r.start)
        o.step()
        dq = float(np.max(np.abs(q[k] - o.get("q")[:s.n])))
        f0, nn, jj, vv = o.get("H_firstnbr"), o.get("H_numnbrs"), o.get("H_jlist"), o.get("H_val")
        Ad = np.diag(1.0 / o.get("Hdia_inv")[:s.n])
        for i in range(s.n):
            np.add.at(Ad[i], jj[f0[i]:f0[i] + nn[i]], vv[f0[i]:f0[i] + nn[i]])
        sd, td = np.linalg.solve(Ad, o.get("b_s")[:s.n]), np.linalg.solve(Ad, o.get("b_t")[:s.n])
        qd = sd - sd.sum() / td.sum() * td
        dq_gpu, dq_orc = float(np.max(np.abs(q[k] - qd))), float(np.max(np.abs(o.get("q")[:s.n] - qd)))
        r = bt.rows(k)
        ob, oe = o.get("bond_index"), o.get("bond_end")
        assert np.array_equal(be[r] - bi[r], oe[:s.n] - ob[:s.n]), k
        gsel = np.concatenate([np.arange(bi[i], be[i]) for i in range(r.start, r.stop)])
        osel = np.concatenate([np.arange(ob[i], oe[i]) for i in range(s.n)])
        dbo = rel_err(BO[gsel], o.get("BO")[osel], 1e-12)
        from .parity import oracle_energy_forces
        eo, fo = oracle_energy_forces(o, np.arange(s.n, dtype=np.int32), s.n)
        eg = dict(zip(["bond", "lp", "ov", "un", "pol", "hb", "ang", "pen", "coa", "tor", "con"], eb[k]))
        de = max(abs(eg[t] - eo[t]) / max(abs(eo[t]), 1.0) for t in eg)
        de_nb = abs(en[k] - (eo["vdw"] + eo["coul"])) / abs(eo["vdw"] + eo["coul"])
        dg = float(np.max(np.abs(grads[k] + fo)) / np.abs(fo).max())    # gradient = -force
        print(f"[c3 oracle] cluster {k} n={s.n}: max|dq| {dq:.2e} (GPU vs dense solve {dq_gpu:.2e}, oracle vs "
              f"dense {dq_orc:.2e}) BO rel {dbo:.2e} energy rel {max(de, de_nb):.2e} gradient rel {dg:.2e}",
              flush=True)
        assert dq <= 1e-8 and dq_gpu <= 1e-8 and dq_orc <= 1e-8 and dbo <= 1e-10
        assert de <= 1e-8 and de_nb <= 1e-8 and dg <= 1e-8, (k, de, de_nb, dg)
