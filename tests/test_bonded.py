"""Bond and atom energies (Merge_Bonds_Atom_Energy_C_New), hydrogen bonds
(Hydrogen_Bonds_C) and the bond-order chain rule of the forces
(Add_All_dBond_to_Forces_C): oracle pins on the CPU and HIP-vs-oracle parity
on the GPU.

Pins (independent of the oracle's code):
  * finite differences of e_bond + e_lp + e_ov + e_un + e_hb (charges do not
    enter them) reproduce the gradient the hydrogen-bond terms and the chain
    rule leave in fCdDelta[:, :3], on water/methanol clusters, compressed
    clusters whose bonds reach the over-coordination branch, and water;
  * the gradient of an isolated cluster sums to zero and has no torque;
  * e_bond equals the published bond-energy expression summed once per bond
    over the oracle's corrected bond orders, with the bond parameters read
    by the independent force-field reader (tests/ffield_kat.py), on a water
    cluster whose bond orders all stay below 1 (the terminal triple-bond
    term, which this force field's gp37 = 2 applies to every bond of order
    >= 1, is then absent; the compressed clusters exercise it);
  * e_pol equals sum 23.02 (chi q + eta/2 q^2) with the reader's chi, eta;
  * e_hb of a water dimer equals the published hydrogen-bond expression
    summed over every H, every O it is bonded to (BO >= 0.01, the oracle's
    bond orders) and every other O within the cutoff, with the reader's
    parameters and the angle and distance from the positions.
The lone-pair, over- and under-coordination expressions are pinned only
through those finite differences and the GPU-vs-oracle parity: no closed
form or reference output covers them.  The C-C lone-pair correction is
never active with this force field (its p_lp3 = gp[5] is 0).
GPU: energies, per-bond coefficients, CdDelta and the gradient of the HIP
path against the oracle on the same bonds, relative 1e-10.
"""
import numpy as np
import pytest

from sw_reaxff_amd import systems

from .ffield_kat import FField

STAGES = ("write_reax_atoms", "write_reax_lists", "estimate_storages", "reset", "write_reax_lists",
          "init_forces", "validate_lists", "bo")


def run_stages(o):
    for st in STAGES:
        rc = o.stage(st)
        assert rc >= 0 if st == "write_reax_lists" else rc == 0, st   # the list pass returns its entry count


def _oracle(oracle_lib, ffield, s, x=None, q=None):
    """Oracle of a non-periodic system through BO (charges fixed)."""
    x = s.x if x is None else x
    o = oracle_lib.Oracle(ffield, s.elements, threads=1)
    o.set_atoms(s.n, np.asarray(x, np.float64), s.type.astype(np.int32), s.tag.astype(np.int32),
                np.arange(s.n, dtype=np.int32), q=q)
    o.build_verlet(12.0)
    run_stages(o)
    return o


def _energy(oracle_lib, ffield, s, x):
    o = _oracle(oracle_lib, ffield, s, x)
    return float(o.atom_energy()[:4].sum()) + o.hydrogen_bonds()


def compressed(s, f):
    """The cluster scaled by f about its centre (f < 1 shortens every bond)."""
    c = s.x.mean(axis=0)
    return systems.System(x=c + f * (s.x - c), v=s.v, type=s.type, tag=s.tag, elements=s.elements, lo=s.lo,
                          hi=s.hi, periodic=False)


def water_cluster(nmol, seed):
    w = systems.water_box(nmol, seed=seed)
    return systems.System(x=w.x, v=w.v, type=w.type, tag=w.tag, elements=w.elements, lo=w.lo - 20.0,
                          hi=w.hi + 20.0, periodic=False)


@pytest.mark.parametrize("case", ["cluster30", "compressed", "water"])
def test_gradient_matches_finite_differences(oracle_lib, ffield, case):
    if case == "cluster30":
        s = systems.cluster(30, seed=5)
    elif case == "compressed":
        s = compressed(systems.cluster(24, seed=8), 0.9)
    else:
        s = water_cluster(8, seed=2)
    o = _oracle(oracle_lib, ffield, s)
    e = o.atom_energy()
    ehb = o.hydrogen_bonds()
    o.dbond_forces()
    g = o.get("fCdDelta").reshape(-1, 4)[: s.n, :3]
    gmax = np.abs(g).max()
    rng = np.random.default_rng(3)
    h = 1e-5
    worst = 0.0
    for a in rng.choice(s.n, size=6, replace=False):
        for k in range(3):
            xp, xm = s.x.copy(), s.x.copy()
            xp[a, k] += h
            xm[a, k] -= h
            fd = (_energy(oracle_lib, ffield, s, xp) - _energy(oracle_lib, ffield, s, xm)) / (2 * h)
            worst = max(worst, abs(fd - g[a, k]) / gmax)
    print(f"[bonded] {case} n={s.n} energies {e} e_hb {ehb:.4e} max|g| {gmax:.3e} worst FD error {worst:.2e}")
    assert worst <= 1e-6


def test_gradient_has_no_net_force_or_torque(oracle_lib, ffield):
    s = systems.cluster(60, seed=11)
    o = _oracle(oracle_lib, ffield, s)
    o.atom_energy()
    assert o.hydrogen_bonds() != 0.0
    o.dbond_forces()
    g = o.get("fCdDelta").reshape(-1, 4)[: s.n, :3]
    scale = np.abs(g).max()
    assert np.abs(g.sum(axis=0)).max() <= 1e-11 * scale * s.n
    torque = np.cross(s.x - s.x.mean(axis=0), g).sum(axis=0)
    assert np.abs(torque).max() <= 1e-10 * scale * s.n


def test_bond_energy_closed_form(oracle_lib, ffield):
    ff = FField(ffield)
    s = water_cluster(12, seed=4)
    o = _oracle(oracle_lib, ffield, s)
    e = o.atom_energy()
    bi, be = o.get("bond_index"), o.get("bond_end")
    bonds, bo = o.get("bonds"), o.get("bo_data")
    BOpi = o.get("BOpi").reshape(-1, 2)
    assert o.get("BO")[: be[s.n - 1]].max() < 1.0     # no triple-bond term in this case
    want = 0.0
    for i in range(s.n):
        for p in range(bi[i], be[i]):
            j = int(bonds["nbr"][p])
            if s.tag[i] > s.tag[j]:
                continue                      # each bond once, by its smaller atom id
            t = ff.pair(s.elements[s.type[i] - 1], s.elements[s.type[j] - 1])
            bos = float(bo["BO_s"][p])
            pw = bos ** t["p_be2"] if bos != 0.0 else 0.0
            want += (-t["De_s"] * bos * np.exp(t["p_be1"] * (1.0 - pw)) - t["De_p"] * BOpi[p, 0]
                     - t["De_pp"] * BOpi[p, 1])
    assert e[0] == pytest.approx(want, rel=1e-12)


def test_polarization_energy_closed_form(oracle_lib, ffield):
    ff = FField(ffield)
    s = systems.cluster(40, seed=6)
    q = np.random.default_rng(1).uniform(-0.8, 0.8, size=s.n)
    o = _oracle(oracle_lib, ffield, s, q=q)
    e = o.atom_energy()
    want = sum(23.02 * (ff.atom(el)["chi"] * qi + ff.eta(el) / 2.0 * qi * qi)
               for el, qi in zip([s.elements[t - 1] for t in s.type], q))
    assert e[4] == pytest.approx(want, rel=1e-12)


def test_hydrogen_bond_closed_form(oracle_lib, ffield):
    """Two water molecules 2.9 A apart (O...O), one H pointing at the other O."""
    ff = FField(ffield)
    ang = np.radians(104.52)
    w = np.array([[0.0, 0.0, 0.0], [0.9572, 0.0, 0.0], [0.9572 * np.cos(ang), 0.9572 * np.sin(ang), 0.0]])
    rot = np.array([[np.cos(0.7), 0.0, np.sin(0.7)], [0.0, 1.0, 0.0], [-np.sin(0.7), 0.0, np.cos(0.7)]])
    x = np.vstack([w, w @ rot.T + np.array([2.9, 0.2, 0.1])])
    s = systems.System(x=x, v=np.zeros_like(x), type=np.array([1, 2, 2, 1, 2, 2], np.int32),
                       tag=np.arange(1, 7, dtype=np.int32), elements=["O", "H"], lo=x.min(0) - 20, hi=x.max(0) + 20,
                       periodic=False)
    o = _oracle(oracle_lib, ffield, s)
    o.atom_energy()
    got = o.hydrogen_bonds()
    bi, be, nbr, BO = o.get("bond_index"), o.get("bond_end"), o.get("bonds")["nbr"], o.get("BO")
    p = ff.hbond("O", "H", "O")
    want, terms = 0.0, 0
    for j in (1, 2, 4, 5):                                  # the H atoms
        for q in range(bi[j], be[j]):
            oi = int(nbr[q])
            if oi not in (0, 3) or BO[q] < 0.01:            # bonds to an O with BO >= 0.01
                continue
            for ok in (0, 3):                               # every other O within the cutoff
                b = x[ok] - x[j]
                r = np.linalg.norm(b)
                if ok == oi or r > 7.5:
                    continue
                a = x[oi] - x[j]
                theta = np.arccos(np.dot(a, b) / (np.linalg.norm(a) * r))
                want += (p["p_hb1"] * (1.0 - np.exp(-p["p_hb2"] * BO[q]))
                         * np.exp(-p["p_hb3"] * (p["r0_hb"] / r + r / p["r0_hb"] - 2.0)) * np.sin(theta / 2.0) ** 4)
                terms += 1
    assert terms >= 4
    assert got == pytest.approx(want, rel=1e-12)
    assert got < -1.0        # a real hydrogen bond


# ---------------------------------------------------------------- GPU parity
def _bond_key(bi, be, nbr, n_rows):
    keys = []
    for i in range(n_rows):
        for p in range(bi[i], be[i]):
            keys.append((i, int(nbr[p]), p))
    keys.sort()
    return np.array([k[2] for k in keys])


def _gpu_vs_oracle(gpu_lib, oracle_lib, ffield, lmp, elements, stype, stag):
    from sw_reaxff_amd import _lib
    h = lmp.system
    _lib.check(h.L.swr_atom_energy(h.h), "swr_atom_energy")
    _lib.check(h.L.swr_hydrogen_bonds(h.h), "swr_hydrogen_bonds")
    h.sync()
    cd_gpu = h.get("fCdDelta").reshape(-1, 4)[:, 3].copy()
    cdbo_gpu = [h.get(k).copy() for k in ("Cdbo", "Cdbopi", "Cdbopi2")]
    ghb_gpu = h.get("fCdDelta").reshape(-1, 4)[:, :3].copy()
    _lib.check(h.L.swr_add_dbond_forces(h.h), "swr_add_dbond_forces")
    h.sync()
    e_gpu = np.append(h.get("bonded_energy"), h.get("hbond_energy"))
    g_gpu = h.get("fCdDelta").reshape(-1, 4)[:, :3]
    n, N = h.scalar("n"), h.scalar("N")
    xa = h.get("x").reshape(-1, 3)[:N]
    owner = h.get("owner")[:N]
    o = oracle_lib.Oracle(ffield, elements, threads=8)
    o.set_atoms(n, xa, stype[owner].astype(np.int32), stag[owner].astype(np.int32), owner.astype(np.int32),
                q=h.get("q")[:N])
    o.build_verlet(12.0)
    run_stages(o)
    e_o = np.append(o.atom_energy(), o.hydrogen_bonds())
    cd_o = o.get("fCdDelta").reshape(-1, 4)[:, 3].copy()
    cdbo_o = [o.get(k).copy() for k in ("Cdbo", "Cdbopi", "Cdbopi2")]
    ghb_o = o.get("fCdDelta").reshape(-1, 4)[:, :3].copy()
    o.dbond_forces()
    g_o = o.get("fCdDelta").reshape(-1, 4)[:, :3]
    print(f"[bonded] n={n} N={N} gpu {e_gpu} oracle {e_o}", flush=True)
    assert np.max(np.abs(ghb_gpu[:N] - ghb_o[:N])) <= 1e-10 * max(np.abs(ghb_o[:N]).max(), 1e-300)
    for k in range(6):
        assert e_gpu[k] == pytest.approx(e_o[k], rel=1e-10, abs=1e-9), k
    # per bond, matched by (atom, partner)
    kg = _bond_key(h.get("bond_index"), h.get("bond_end"), h.get("bonds")["nbr"], N)
    ko = _bond_key(o.get("bond_index"), o.get("bond_end"), o.get("bonds")["nbr"], N)
    assert kg.size == ko.size
    for a, b in zip(cdbo_gpu, cdbo_o):
        scale = np.abs(b[ko]).max()
        assert np.max(np.abs(a[kg] - b[ko])) <= 1e-10 * scale
    assert np.max(np.abs(cd_gpu[:N] - cd_o[:N])) <= 1e-10 * np.abs(cd_o[:N]).max()
    assert np.max(np.abs(g_gpu[:N] - g_o[:N])) <= 1e-10 * np.abs(g_o[:N]).max()


@pytest.mark.gpu
def test_gpu_bonded_water_box(gpu_lib, oracle_lib, ffield):
    from .parity import harness
    s = systems.water_box(648, seed=3)
    lmp = harness(s, ffield, tol=1e-10)
    _gpu_vs_oracle(gpu_lib, oracle_lib, ffield, lmp, s.elements, s.type, s.tag)


@pytest.mark.gpu
@pytest.mark.parametrize("f", [1.0, 0.9])
def test_gpu_bonded_cluster(gpu_lib, oracle_lib, ffield, f):
    from .parity import harness
    s = compressed(systems.cluster(700, seed=9), f)
    lmp = harness(s, ffield, tol=1e-10)
    _gpu_vs_oracle(gpu_lib, oracle_lib, ffield, lmp, s.elements, s.type, s.tag)
