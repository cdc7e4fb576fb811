"""Valence angles (the valence part of Merge_Torsion_Valence_Angles).

Oracle pins (CPU):
  * finite differences of e_ang + e_pen + e_coa reproduce the gradient the
    oracle's valence pass plus the bond-order chain rule give (water/methanol
    clusters, a compressed one, nitromethane beside propane for the penalty
    and coalition terms, water);
  * zero net force and torque;
  * an isolated water molecule: e_ang equals the published angle expression
    f7(BO_a) f7(BO_b) f8(Delta_j) p_val1 (1 - exp(-p_val2 (theta_0 -
    theta)^2)) summed over its three triples, with theta from the positions,
    the parameters read from the raw ffield text and theta_0 from SBO; these
    triples have no penalty or coalition parameters, so those energies
    vanish.
GPU parity: energies, Cdbo/Cdbopi/Cdbopi2 per bond, CdDelta and the
gradient after the chain rule at 1e-10 relative (water box, clusters).

Torsions (the torsion part of the same function): finite differences of
e_tor + e_con (the same four systems), zero net force and torque, the
dihedral dependence of e_tor and e_con on an isolated chain against the
published functional form (cosine series, zero at trans, V1 : V3 from the
raw ffield text), and the same GPU parity.
"""
import math

import numpy as np
import pytest

from sw_reaxff_amd import systems

from .ffield_kat import FField
from .test_bonded import _bond_key, compressed, run_stages, water_cluster


def _oracle(oracle_lib, ffield, s, x=None):
    x = s.x if x is None else x
    o = oracle_lib.Oracle(ffield, s.elements, threads=1)
    o.set_atoms(s.n, np.asarray(x, np.float64), s.type.astype(np.int32), s.tag.astype(np.int32),
                np.arange(s.n, dtype=np.int32))
    o.build_verlet(12.0)
    run_stages(o)
    return o


def _energy(oracle_lib, ffield, s, x):
    return float(_oracle(oracle_lib, ffield, s, x).valence_angles().sum())


def nitromethane_propane(seed):
    """Nitromethane (O-N-O: penalty and coalition terms) beside a propane
    carbon chain (C-C-C: penalty), slightly distorted."""
    els = ["C", "H", "O", "N"]
    x = np.array([[0.0, 0.0, 0.0], [1.49, 0.0, 0.0], [2.10, 1.07, 0.0], [2.10, -1.07, 0.0],
                  [-0.36, 1.03, 0.0], [-0.36, -0.51, 0.89], [-0.36, -0.51, -0.89],
                  [0.0, 4.5, 0.0], [1.53, 4.5, 0.0], [2.04, 5.94, 0.0],
                  [-0.37, 3.48, 0.0], [-0.37, 5.0, 0.89], [1.90, 3.98, 0.89], [3.14, 5.94, 0.0]])
    t = np.array([1, 4, 3, 3, 2, 2, 2, 1, 1, 1, 2, 2, 2, 2])
    x = x + np.random.default_rng(seed).normal(scale=0.05, size=x.shape) + 10.0
    return systems.System(x=x, v=np.zeros_like(x), type=t, tag=np.arange(1, len(t) + 1), elements=els,
                          lo=np.zeros(3), hi=np.full(3, 30.0), periodic=False)


@pytest.mark.parametrize("case", ["cluster30", "compressed", "nitro", "water"])
def test_valence_gradient_matches_finite_differences(oracle_lib, ffield, case):
    if case == "cluster30":
        s = systems.cluster(30, seed=5)
    elif case == "compressed":
        s = compressed(systems.cluster(24, seed=8), 0.9)
    elif case == "nitro":
        s = nitromethane_propane(3)
    else:
        s = water_cluster(8, seed=2)
    o = _oracle(oracle_lib, ffield, s)
    e = o.valence_angles()
    o.dbond_forces()
    g = o.get("fCdDelta").reshape(-1, 4)[: s.n, :3]
    gmax = np.abs(g).max()
    assert e[0] != 0.0
    if case == "nitro":
        assert e[1] != 0.0 and e[2] != 0.0
    rng = np.random.default_rng(7)
    h = 1e-5
    worst = 0.0
    for a in rng.choice(s.n, size=6, replace=False):
        for k in range(3):
            xp, xm = s.x.copy(), s.x.copy()
            xp[a, k] += h
            xm[a, k] -= h
            fd = (_energy(oracle_lib, ffield, s, xp) - _energy(oracle_lib, ffield, s, xm)) / (2 * h)
            worst = max(worst, abs(fd - g[a, k]) / gmax)
    print(f"[valence] {case} n={s.n} e_ang/e_pen/e_coa {e} max|g| {gmax:.3e} worst FD error {worst:.2e}")
    assert worst <= 1e-6


def test_valence_gradient_has_no_net_force_or_torque(oracle_lib, ffield):
    s = systems.cluster(60, seed=11)
    o = _oracle(oracle_lib, ffield, s)
    o.valence_angles()
    o.dbond_forces()
    g = o.get("fCdDelta").reshape(-1, 4)[: s.n, :3]
    scale = np.abs(g).max()
    assert np.abs(g.sum(axis=0)).max() <= 1e-11 * scale * s.n
    torque = np.cross(s.x - s.x.mean(axis=0), g).sum(axis=0)
    assert np.abs(torque).max() <= 1e-10 * scale * s.n


def test_water_molecule_angle_energy_closed_form(oracle_lib, ffield):
    """Every triple of an isolated water molecule: H-O-H centred on O and,
    through the weak H-H bond (BO ~ 0.012 > thb_cut), O-H-H centred on each
    H, each f7(BO_a) f7(BO_b) f8(Delta_boc) p_val1 (1 - exp(-p_val2
    (theta_0 - theta)^2)) with theta from the positions."""
    ff = FField(ffield)
    els = ["H", "O"]
    x = systems.water_body()[[0, 1, 2]] + 5.0
    s = systems.System(x=x, v=np.zeros_like(x), type=np.array([2, 1, 1]), tag=np.array([1, 2, 3]),
                       elements=els, lo=np.zeros(3), hi=np.full(3, 20.0), periodic=False)
    o = _oracle(oracle_lib, ffield, s)
    e = o.valence_angles()
    bi, be = o.get("bond_index"), o.get("bond_end")
    nbr = o.get("bonds")["nbr"]
    BO = o.get("BO")
    BOpi = o.get("BOpi").reshape(-1, 2)
    Dboc_all = o.get("bo_dboc").reshape(-1, 2)[:, 1]
    vlpex, nlp = o.get("vlpex"), o.get("nlp")
    gp = ff.gp
    thb_cut, thb_cutsq = 0.001, 0.00001
    pi_reference = 3.14159265                   # the reference's constPI
    el = ["O", "H", "H"]
    want, ntriples = 0.0, 0
    for j in range(3):
        a = ff.atom(el[j])
        p_val3, p_val5 = a["p_val3"], a["p_val5"]
        slots = list(range(bi[j], be[j]))
        SBOp = sum(BOpi[p].sum() for p in slots)
        prod = math.prod(math.exp(-BO[p] ** 8) for p in slots)
        vlpadj = 0.0 if vlpex[j] >= 0 else nlp[j]
        SBO = SBOp + (1 - prod) * (-Dboc_all[j] - gp[33] * vlpadj)
        if SBO <= 0:
            SBO2 = 0.0
        elif SBO <= 1:
            SBO2 = SBO ** gp[16]
        elif SBO < 2:
            SBO2 = 2 - (2 - SBO) ** gp[16]
        else:
            SBO2 = 2.0
        for u in range(len(slots)):
            for w in range(u + 1, len(slots)):
                pa, pb = slots[u], slots[w]
                if not (BO[pa] > thb_cut and BO[pb] > thb_cut and BO[pa] * BO[pb] > thb_cutsq):
                    continue
                ka, kb = int(nbr[pa]), int(nbr[pb])
                v1, v2 = x[ka] - x[j], x[kb] - x[j]
                theta = math.acos(v1 @ v2 / (np.linalg.norm(v1) * np.linalg.norm(v2)))
                for prm in ff.angle(el[ka], el[j], el[kb]):
                    if abs(prm["p_val1"]) <= 0.001:
                        continue
                    ntriples += 1
                    theta0 = (180.0 - prm["theta_00"] * (1 - math.exp(-gp[17] * (2 - SBO2)))) * pi_reference / 180.0
                    f7a = 1 - math.exp(-p_val3 * (BO[pa] - thb_cut) ** prm["p_val4"])
                    f7b = 1 - math.exp(-p_val3 * (BO[pb] - thb_cut) ** prm["p_val4"])
                    e6, e7 = math.exp(gp[14] * Dboc_all[j]), math.exp(-prm["p_val7"] * Dboc_all[j])
                    f8 = p_val5 - (p_val5 - 1) * (2 + e6) / (1 + e6 + e7)
                    want += f7a * f7b * f8 * prm["p_val1"] * (1 - math.exp(-prm["p_val2"] * (theta0 - theta) ** 2))
    print(f"[valence] water molecule: {ntriples} triples, e_ang {e[0]:.10e} closed form {want:.10e}")
    assert ntriples == 3
    assert e[0] == pytest.approx(want, rel=1e-12)
    assert e[1] == 0.0 and e[2] == 0.0      # no penalty or coalition parameters for these triples


# ---------------------------------------------------------------- GPU parity
def _gpu_vs_oracle(oracle_lib, ffield, lmp, elements, stype, stag):
    from sw_reaxff_amd import _lib
    h = lmp.system
    _lib.check(h.L.swr_valence_angles(h.h), "swr_valence_angles")
    h.sync()
    cd_gpu = h.get("fCdDelta").reshape(-1, 4)[:, 3].copy()
    cdbo_gpu = [h.get(k).copy() for k in ("Cdbo", "Cdbopi", "Cdbopi2")]
    gv_gpu = h.get("fCdDelta").reshape(-1, 4)[:, :3].copy()
    _lib.check(h.L.swr_add_dbond_forces(h.h), "swr_add_dbond_forces")
    h.sync()
    e_gpu = h.get("valence_energy")
    ea = h.get("valence_eatom")
    g_gpu = h.get("fCdDelta").reshape(-1, 4)[:, :3]
    n, N = h.scalar("n"), h.scalar("N")
    xa = h.get("x").reshape(-1, 3)[:N]
    owner = h.get("owner")[:N]
    o = oracle_lib.Oracle(ffield, elements, threads=8)
    o.set_atoms(n, xa, stype[owner].astype(np.int32), stag[owner].astype(np.int32), owner.astype(np.int32),
                q=h.get("q")[:N])
    o.build_verlet(12.0)
    run_stages(o)
    e_o = o.valence_angles()
    cd_o = o.get("fCdDelta").reshape(-1, 4)[:, 3].copy()
    cdbo_o = [o.get(k).copy() for k in ("Cdbo", "Cdbopi", "Cdbopi2")]
    gv_o = o.get("fCdDelta").reshape(-1, 4)[:, :3].copy()
    o.dbond_forces()
    g_o = o.get("fCdDelta").reshape(-1, 4)[:, :3]
    print(f"[valence] n={n} N={N} gpu {e_gpu} oracle {e_o}", flush=True)
    for k in range(3):
        assert e_gpu[k] == pytest.approx(e_o[k], rel=1e-10, abs=1e-9), k
    assert ea.reshape(3, -1)[:, :n].sum(axis=1) == pytest.approx(e_gpu, rel=1e-12, abs=1e-12)
    kg = _bond_key(h.get("bond_index"), h.get("bond_end"), h.get("bonds")["nbr"], N)
    ko = _bond_key(o.get("bond_index"), o.get("bond_end"), o.get("bonds")["nbr"], N)
    assert kg.size == ko.size
    for a, b in zip(cdbo_gpu, cdbo_o):
        scale = np.abs(b[ko]).max()
        assert np.max(np.abs(a[kg] - b[ko])) <= 1e-10 * scale
    assert np.max(np.abs(cd_gpu[:N] - cd_o[:N])) <= 1e-10 * np.abs(cd_o[:N]).max()
    assert np.max(np.abs(gv_gpu[:N] - gv_o[:N])) <= 1e-10 * np.abs(gv_o[:N]).max()
    assert np.max(np.abs(g_gpu[:N] - g_o[:N])) <= 1e-10 * np.abs(g_o[:N]).max()


@pytest.mark.gpu
def test_gpu_valence_water_box(gpu_lib, oracle_lib, ffield):
    from .parity import harness
    s = systems.water_box(648, seed=3)
    lmp = harness(s, ffield, tol=1e-10)
    _gpu_vs_oracle(oracle_lib, ffield, lmp, s.elements, s.type, s.tag)


@pytest.mark.gpu
@pytest.mark.parametrize("f", [1.0, 0.9])
def test_gpu_valence_cluster(gpu_lib, oracle_lib, ffield, f):
    from .parity import harness
    s = compressed(systems.cluster(700, seed=9), f)
    lmp = harness(s, ffield, tol=1e-10)
    _gpu_vs_oracle(oracle_lib, ffield, lmp, s.elements, s.type, s.tag)


# ---------------------------------------------------------------- torsions
def _tor_energy(oracle_lib, ffield, s, x):
    return float(_oracle(oracle_lib, ffield, s, x).torsion_angles().sum())


@pytest.mark.parametrize("case", ["cluster30", "compressed", "nitro", "water"])
def test_torsion_gradient_matches_finite_differences(oracle_lib, ffield, case):
    """Torsion part of Merge_Torsion_Valence_Angles: finite differences of
    e_tor + e_con against the gradient of the oracle's torsion pass plus the
    bond-order chain rule."""
    if case == "cluster30":
        s = systems.cluster(30, seed=5)
    elif case == "compressed":
        s = compressed(systems.cluster(24, seed=8), 0.9)
    elif case == "nitro":
        s = nitromethane_propane(3)
    else:
        s = water_cluster(8, seed=2)
    o = _oracle(oracle_lib, ffield, s)
    e = o.torsion_angles()
    o.dbond_forces()
    g = o.get("fCdDelta").reshape(-1, 4)[: s.n, :3]
    gmax = np.abs(g).max()
    assert e[0] != 0.0
    if case != "water":                   # water's quadruples have p_cot1 = 0
        assert e[1] != 0.0
    rng = np.random.default_rng(9)
    h = 1e-5
    worst = 0.0
    for a in rng.choice(s.n, size=6, replace=False):
        for k in range(3):
            xp, xm = s.x.copy(), s.x.copy()
            xp[a, k] += h
            xm[a, k] -= h
            fd = (_tor_energy(oracle_lib, ffield, s, xp) - _tor_energy(oracle_lib, ffield, s, xm)) / (2 * h)
            worst = max(worst, abs(fd - g[a, k]) / gmax)
    print(f"[torsion] {case} n={s.n} e_tor/e_con {e} max|g| {gmax:.3e} worst FD error {worst:.2e}")
    assert worst <= 1e-6


def test_torsion_gradient_has_no_net_force_or_torque(oracle_lib, ffield):
    s = systems.cluster(60, seed=11)
    o = _oracle(oracle_lib, ffield, s)
    assert o.torsion_angles()[0] != 0.0
    o.dbond_forces()
    g = o.get("fCdDelta").reshape(-1, 4)[: s.n, :3]
    scale = np.abs(g).max()
    assert np.abs(g.sum(axis=0)).max() <= 1e-11 * scale * s.n
    torque = np.cross(s.x - s.x.mean(axis=0), g).sum(axis=0)
    assert np.abs(torque).max() <= 1e-10 * scale * s.n


def _chain(els4, r, theta, omega):
    """Atoms i-j-k-l, bonds r, both bond angles theta, dihedral omega
    (0 = cis: i and l on the same side of the j-k bond)."""
    j = np.zeros(3)
    k = np.array([r[1], 0.0, 0.0])
    i = j + r[0] * np.array([-math.cos(math.pi - theta), math.sin(math.pi - theta), 0.0])
    l = k + r[2] * np.array([math.cos(math.pi - theta), math.sin(math.pi - theta) * math.cos(omega),
                             math.sin(math.pi - theta) * math.sin(omega)])
    x = np.array([i, j, k, l]) + 10.0
    order = ["C", "H", "O", "N"]
    t = np.array([order.index(e) + 1 for e in els4])
    return systems.System(x=x, v=np.zeros_like(x), type=t, tag=np.arange(1, 5), elements=order,
                          lo=np.zeros(3), hi=np.full(3, 30.0), periodic=False)


@pytest.mark.parametrize("els4", [("N", "C", "N", "N"), ("O", "C", "O", "O")])
def test_torsion_energy_dihedral_series(oracle_lib, ffield, els4):
    """The torsion and conjugation energies' dependence on the dihedral
    angle, against the published functional form.  Rotating l about the j-k
    bond of an isolated chain keeps every bond length and bond angle (so
    every bond order, f10, f11, f12 and sin(theta)); then
      E_tor(w) = f10 s [V1 (1 + cos w) + V2 e (1 - cos 2w) + V3 (1 + cos 3w)] / 2,
      E_con(w) = p_cot1 f12 [1 + (cos^2 w - 1) s],   s = sin th_ijk sin th_jkl,
    with 0 < e = exp(p_tor1 (2 - BOpi_jk - f11)^2) <= 1 (p_tor1 < 0).
    Fitted over twelve dihedrals: E_tor is a cosine series of degree three
    with no sine terms, vanishes at trans, its cos w and cos 3w amplitudes
    stand as V1 : V3 from the raw ffield text, its constant is the sum of
    those two less the cos 2w one, and its cos 2w amplitude carries -V2 e;
    E_con is c0 + c2 cos 2w with c0 / c2 = 2 / s - 1.  A dropped term, a
    wrong multiple of w or a sign slip in any term fails one of these."""
    ff = FField(ffield)
    prm = ff.torsion(*els4)
    assert prm is not None and prm["V1"] != 0 and prm["V2"] != 0 and prm["V3"] != 0 and prm["p_cot1"] != 0
    # long bonds and wide angles put the 1-3 pairs at 3.55 A, beyond bo_cut
    theta = math.radians(160.0)
    r = (1.8, 1.8, 1.8)
    ws = np.radians(np.arange(0.0, 360.0, 30.0))
    et, ec = [], []
    for w in ws:
        s = _chain(els4, r, theta, w)
        o = _oracle(oracle_lib, ffield, s)
        bi, be, nbr = o.get("bond_index"), o.get("bond_end"), o.get("bonds")["nbr"]
        pairs = {tuple(sorted((a, int(nbr[p])))) for a in range(4) for p in range(bi[a], be[a])}
        assert pairs == {(0, 1), (1, 2), (2, 3)}, (np.degrees(w), pairs)   # the chain's bonds only
        e = o.torsion_angles()
        et.append(e[0])
        ec.append(e[1])
    et, ec = np.array(et), np.array(ec)
    A = np.stack([np.ones_like(ws), np.cos(ws), np.cos(2 * ws), np.cos(3 * ws), np.sin(ws), np.sin(2 * ws),
                  np.sin(3 * ws)], axis=1)
    a, *_ = np.linalg.lstsq(A, et, rcond=None)
    c, *_ = np.linalg.lstsq(A, ec, rcond=None)
    st, sc = np.abs(et).max(), np.abs(ec).max()
    print(f"[torsion series] {els4}: E_tor fit {a} E_con fit {c}", flush=True)
    assert st > 0 and sc > 0
    assert np.abs(A @ a - et).max() <= 1e-10 * st and np.abs(a[4:]).max() <= 1e-10 * st
    assert abs(et[6]) <= 1e-12 * st                                  # trans (180 degrees)
    assert a[1] / a[3] == pytest.approx(prm["V1"] / prm["V3"], rel=1e-9)
    assert a[0] == pytest.approx(a[1] - a[2] + a[3], rel=1e-9, abs=1e-12 * st)
    ev = -a[2] / a[1] * prm["V1"] / prm["V2"]                        # e of the V2 term
    assert 0.0 < ev <= 1.0
    sth = math.sin(theta) ** 2
    assert np.abs(A @ c - ec).max() <= 1e-10 * sc and np.abs(c[[1, 3, 4, 5, 6]]).max() <= 1e-10 * sc
    assert c[0] / c[2] == pytest.approx(2.0 / sth - 1.0, rel=1e-9)


def _torsion_fix_floor(bi, be, nbr, N):
    """Bounds on the fixed-point rounding in the torsion kernel's outputs.

    The terms of the atoms i, k and l of a quadruple and the Cdbo of its k-l
    and j-i slots are added into 64-bit accumulators in units of 2^-40, so
    each add rounds by at most 2^-41.  The number of adds an entry receives is
    bounded from the bond lists (every listed bond counted, whatever its bond
    order): atom a takes one add per (j, k, i) with l = a, one per (j, k)
    with i = a and one per bond j-k with k = a; the slot of bond u-v in u's
    list one per (j, i) around it as the k-l slot and one per k as the j-i
    slot.  Returns (per atom, per slot) bounds."""
    bi, be = bi[:N].astype(np.int64), be[:N].astype(np.int64)
    deg = be - bi
    slots = np.concatenate([np.arange(bi[i], be[i]) for i in range(N)]) if N else np.zeros(0, np.int64)
    row = np.repeat(np.arange(N), deg)
    nb = nbr[slots].astype(np.int64)
    S = np.zeros(N, np.int64)                    # S(u) = sum over u's neighbours w of deg(w) - 1
    np.add.at(S, row, deg[nb] - 1)
    as_kl = S[row] - (deg[nb] - 1)               # (j, i) pairs around the k-l slot u -> v
    per_atom = np.zeros(N, np.int64)
    np.add.at(per_atom, nb, as_kl + (deg[row] - 1) + 1)
    per_slot = np.zeros(int(be.max()) if N else 0, np.int64)
    per_slot[slots] = as_kl + (deg[row] - 1)
    return per_atom * 2.0 ** -41, per_slot * 2.0 ** -41


def _gpu_tor_vs_oracle(oracle_lib, ffield, lmp, elements, stype, stag):
    from sw_reaxff_amd import _lib
    h = lmp.system
    _lib.check(h.L.swr_torsion_angles(h.h), "swr_torsion_angles")
    h.sync()
    cd_gpu = h.get("fCdDelta").reshape(-1, 4)[:, 3].copy()
    cdbo_gpu = [h.get(k).copy() for k in ("Cdbo", "Cdbopi")]
    gt_gpu = h.get("fCdDelta").reshape(-1, 4)[:, :3].copy()
    _lib.check(h.L.swr_add_dbond_forces(h.h), "swr_add_dbond_forces")
    h.sync()
    e_gpu = h.get("torsion_energy")
    ea = h.get("torsion_eatom")
    g_gpu = h.get("fCdDelta").reshape(-1, 4)[:, :3]
    n, N = h.scalar("n"), h.scalar("N")
    xa = h.get("x").reshape(-1, 3)[:N]
    owner = h.get("owner")[:N]
    o = oracle_lib.Oracle(ffield, elements, threads=8)
    o.set_atoms(n, xa, stype[owner].astype(np.int32), stag[owner].astype(np.int32), owner.astype(np.int32),
                q=h.get("q")[:N])
    o.build_verlet(12.0)
    run_stages(o)
    e_o = o.torsion_angles()
    cd_o = o.get("fCdDelta").reshape(-1, 4)[:, 3].copy()
    cdbo_o = [o.get(k).copy() for k in ("Cdbo", "Cdbopi")]
    gt_o = o.get("fCdDelta").reshape(-1, 4)[:, :3].copy()
    o.dbond_forces()
    g_o = o.get("fCdDelta").reshape(-1, 4)[:, :3]
    print(f"[torsion] n={n} N={N} gpu {e_gpu} oracle {e_o}", flush=True)
    for k in range(2):
        assert e_gpu[k] == pytest.approx(e_o[k], rel=1e-10, abs=1e-9), k
    assert ea.reshape(2, -1)[:, :n].sum(axis=1) == pytest.approx(e_gpu, rel=1e-12, abs=1e-12)
    kg = _bond_key(h.get("bond_index"), h.get("bond_end"), h.get("bonds")["nbr"], N)
    ko = _bond_key(o.get("bond_index"), o.get("bond_end"), o.get("bonds")["nbr"], N)
    assert kg.size == ko.size
    # the terms of atoms other than the centre and the Cdbo of the k-l and
    # j-i slots are summed in 64-bit fixed point: beside the 1e-10 relative
    # bar, each entry may differ by its number of adds times 2^-41
    fa, fs = _torsion_fix_floor(h.get("bond_index"), h.get("bond_end"), h.get("bonds")["nbr"], N)
    print(f"[torsion] fixed-point floors: atoms <= {fa.max():.2e}, slots <= {fs.max():.2e}", flush=True)
    scale = np.abs(cdbo_o[0][ko]).max()
    assert np.all(np.abs(cdbo_gpu[0][kg] - cdbo_o[0][ko]) <= 1e-10 * scale + fs[kg])
    scale = np.abs(cdbo_o[1][ko]).max()                   # Cdbopi: written directly, no fixed point
    assert np.max(np.abs(cdbo_gpu[1][kg] - cdbo_o[1][ko])) <= 1e-10 * scale
    assert np.all(np.abs(cd_gpu[:N] - cd_o[:N]) <= 1e-10 * np.abs(cd_o[:N]).max() + fa)
    assert np.all(np.abs(gt_gpu[:N] - gt_o[:N]) <= 1e-10 * np.abs(gt_o[:N]).max() + fa[:, None])
    assert np.max(np.abs(g_gpu[:N] - g_o[:N])) <= 1e-10 * np.abs(g_o[:N]).max()


@pytest.mark.gpu
def test_gpu_torsion_water_box(gpu_lib, oracle_lib, ffield):
    from .parity import harness
    s = systems.water_box(648, seed=3)
    lmp = harness(s, ffield, tol=1e-10)
    _gpu_tor_vs_oracle(oracle_lib, ffield, lmp, s.elements, s.type, s.tag)


@pytest.mark.gpu
@pytest.mark.parametrize("f", [1.0, 0.9])
def test_gpu_torsion_cluster(gpu_lib, oracle_lib, ffield, f):
    from .parity import harness
    s = compressed(systems.cluster(700, seed=9), f)
    lmp = harness(s, ffield, tol=1e-10)
    _gpu_tor_vs_oracle(oracle_lib, ffield, lmp, s.elements, s.type, s.tag)


def test_torsion_fix_floor_counts_on_a_chain():
    """A four-atom chain 0-1-2-3: the bound counts both orientations of the
    one quadruple (atom 3 as l, as i and as k), three adds an end atom."""
    bi = np.array([0, 1, 3, 5])
    be = np.array([1, 3, 5, 6])
    nbr = np.array([1, 0, 2, 1, 3, 2])
    fa, fs = _torsion_fix_floor(bi, be, nbr, 4)
    assert np.array_equal(fa / 2.0 ** -41, [3, 3, 3, 3])
    # slot 0 (0->1): as k-l slot, the (j, i) pairs around 0 other than 1: none; as j-i slot: deg(0) - 1 = 0
    # slot 1 (1->0): j = 2 with its i = 3 (one pair); as j-i slot: deg(1) - 1 = 1
    assert np.array_equal(fs / 2.0 ** -41, [0, 2, 1, 1, 2, 0])


# ---------------------------------------------------------------- near-linear torsions
def near_linear_chain(dev):
    """Four carbons C0-C1-C2-C3 1.40 A apart: the C0-C1-C2 angle within dev
    radians of linear, C1-C2-C3 at 110 degrees and out of the plane, so the
    dihedral is ill-defined as dev goes to zero.  The reference clamps the
    sine at MIN_SINE (1e-10) and returns finite values."""
    b, th = 1.40, math.radians(110.0)
    x = np.zeros((4, 3))
    x[2] = [b, 0.0, 0.0]
    x[0] = [-b * math.cos(dev), b * math.sin(dev), 0.0]
    x[3] = [b - b * math.cos(th), b * math.sin(th) * math.cos(0.7), b * math.sin(th) * math.sin(0.7)]
    x = x + 10.0
    return systems.System(x=x, v=np.zeros_like(x), type=np.array([1, 1, 1, 1]), tag=np.arange(1, 5),
                          elements=["C", "H", "O", "N"], lo=np.zeros(3), hi=np.full(3, 30.0), periodic=False)


def _oracle_torsion(oracle_lib, ffield, s, x, q):
    o = oracle_lib.Oracle(ffield, s.elements, threads=1)
    o.set_atoms(s.n, x, s.type.astype(np.int32), s.tag.astype(np.int32), np.arange(s.n, dtype=np.int32), q=q)
    o.build_verlet(12.0)
    run_stages(o)
    e = o.torsion_angles()
    cd = o.get("fCdDelta").reshape(-1, 4)[: s.n].copy()
    o.dbond_forces()
    return e, cd, o.get("fCdDelta").reshape(-1, 4)[: s.n, :3].copy()


@pytest.mark.gpu
@pytest.mark.parametrize("dev", [1e-3, 1e-4, 1e-6, 1e-8, 0.0])
def test_gpu_torsion_near_linear(gpu_lib, oracle_lib, ffield, dev):
    """A torsion with one angle near or at linear is a valid geometry: the
    GPU returns finite values (no SWR_ERR_NUMERIC), and its energies,
    CdDelta and forces equal the oracle's within 1e-10 relative plus ten
    times the oracle's own sensitivity to its input (the spread of its
    results over positions changed by a few ulps, which grows as dev^-2:
    4e-9 at 1e-3, 4e-3 at 1e-6).  Below ~1e-7 rad the oracle's forces are
    rounding noise (2.5e3 at 1e-8 where the limit is 12.6) and that bar is
    loose; the GPU, evaluating the same expressions without contraction,
    lands within ~1e-12 of the oracle's value there too."""
    from sw_reaxff_amd import _lib
    from .parity import harness
    s = near_linear_chain(dev)
    lmp = harness(s, ffield, tol=1e-10)
    h = lmp.system
    _lib.check(h.L.swr_torsion_angles(h.h), "swr_torsion_angles")
    h.sync()
    cd_g = h.get("fCdDelta").reshape(-1, 4)[: s.n].copy()
    _lib.check(h.L.swr_add_dbond_forces(h.h), "swr_add_dbond_forces")
    h.sync()                                              # raises on SWR_ERR_NUMERIC
    e_g = h.get("torsion_energy").copy()
    g_g = h.get("fCdDelta").reshape(-1, 4)[: s.n, :3].copy()
    assert np.all(np.isfinite(cd_g)) and np.all(np.isfinite(g_g)) and np.all(np.isfinite(e_g))
    x = h.get("x").reshape(-1, 3)[: s.n].copy()
    q = h.get("q")[: s.n].copy()
    e_o, cd_o, g_o = _oracle_torsion(oracle_lib, ffield, s, x, q)
    rng = np.random.default_rng(4)
    spread, spread_e = 0.0, np.zeros(2)
    for _ in range(8):
        xp = x * (1.0 + rng.integers(-2, 3, size=x.shape) * 2.0 ** -52)
        ep, cdp, gp = _oracle_torsion(oracle_lib, ffield, s, xp, q)
        spread = max(spread, np.abs(cdp - cd_o).max(), np.abs(gp - g_o).max())
        spread_e = np.maximum(spread_e, np.abs(ep - e_o))
    print(f"[torsion] dev {dev:g}: e gpu {e_g} oracle {e_o} spread {spread_e}; |g| {np.abs(g_o).max():.3e} "
          f"gpu-oracle {np.abs(g_g - g_o).max():.2e} oracle spread {spread:.2e}", flush=True)
    assert np.all(np.abs(e_g - e_o) <= 1e-10 * np.abs(e_o) + 10.0 * spread_e + 1e-14)
    scale = np.abs(g_o).max()
    assert np.abs(cd_g - cd_o).max() <= 1e-10 * scale + 10.0 * spread
    assert np.abs(g_g - g_o).max() <= 1e-10 * scale + 10.0 * spread
