"""Non-bonded pass (vdW_Coulomb_Energy_Full_C): oracle pins on the CPU and
HIP-vs-oracle parity on the GPU.

Pins (independent of the oracle's code):
  * two atoms: the energy equals the closed form computed from the raw
    force-field text (tests/ffield_kat.py: mixing rules, off-diagonal
    overrides, 7th-order taper, shielded Morse, shielded Coulomb) and the
    gradient equals the closed form's derivative along the bond;
  * clusters: finite differences of the total energy (charges fixed)
    reproduce the gradient; the gradients sum to zero (every pair appears in
    both atoms' full lists);
  * the inner-wall form (vdw_type 3) through a force field whose atoms all
    get inner-wall parameters (the repository's force field is vdw_type 1).
GPU: energies, virial and fCdDelta of the HIP pass against the oracle on
the same far lists and charges, relative 1e-10 (the kernel forms the powers
from exp/log and the cube root with cbrt, and sums in another order).
"""
import numpy as np
import pytest

from sw_reaxff_amd import systems

from .ffield_kat import FField


def _oracle(oracle_lib, ffield, elements, x, types, q, owner=None, n=None):
    """Oracle with fixed charges: packed atoms, far lists, capacities, Reset."""
    x = np.asarray(x, dtype=np.float64)
    N = x.shape[0]
    n = N if n is None else n
    owner = np.arange(N, dtype=np.int32) if owner is None else np.asarray(owner, np.int32)
    o = oracle_lib.Oracle(ffield, elements, threads=4)
    o.set_atoms(n, x, np.asarray(types, np.int32), owner + 1, owner, q=np.asarray(q, dtype=np.float64))
    o.build_verlet(12.0)
    for st in ("write_reax_atoms", "write_reax_lists", "estimate_storages", "reset"):
        o.stage(st)
    return o


def _energy_grad(oracle_lib, ffield, elements, x, types, q):
    o = _oracle(oracle_lib, ffield, elements, x, types, q)
    ev, ec, vir = o.vdw_coulomb()
    g = o.get("fCdDelta").reshape(-1, 4)[:, :3].copy()
    return ev, ec, vir, g


@pytest.mark.parametrize("a,b", [("H", "O"), ("O", "O"), ("C", "H"), ("N", "O")])
@pytest.mark.parametrize("r", [1.4, 3.1, 6.5, 9.9])
def test_two_atom_closed_form(oracle_lib, ffield, a, b, r):
    ff = FField(ffield)
    els = ["C", "H", "O", "N"]
    qa, qb = 0.31, -0.47
    x = np.array([[0.0, 0.0, 0.0], [r * 0.6, r * 0.8, 0.0]])
    ev, ec, vir, g = _energy_grad(oracle_lib, ffield, els, x, [els.index(a) + 1, els.index(b) + 1], [qa, qb])
    d = float(np.sqrt(x[1, 0] * x[1, 0] + x[1, 1] * x[1, 1]))
    Ev, Ec = ff.nonbonded_pair_energy(a, b, d, qa, qb)
    # near the cutoff the taper itself is ~1e-7: its Horner form and the
    # closed form then differ by rounding of order 1e-13 absolute
    assert ev == pytest.approx(Ev, rel=1e-12, abs=1e-12)
    assert ec == pytest.approx(Ec, rel=1e-12, abs=1e-12)
    h = 1e-4       # central difference: truncation ~h^2, rounding ~1e-13/h
    dEdr = (sum(ff.nonbonded_pair_energy(a, b, d + h, qa, qb)) - sum(ff.nonbonded_pair_energy(a, b, d - h, qa, qb))) / (2 * h)
    u = (x[0] - x[1]) / d                          # d r / d x_0
    assert np.allclose(g[0], dEdr * u, rtol=1e-6, atol=1e-9 * max(1.0, abs(dEdr)))
    assert np.allclose(g[1], -dEdr * u, rtol=1e-6, atol=1e-9 * max(1.0, abs(dEdr)))
    # virial: half of each of the two full-list entries, fpair del del, where
    # atom 0's gradient is fpair (x1 - x0)
    dv = x[1] - x[0]
    assert vir[0] == pytest.approx(g[0][0] * dv[0], rel=1e-12, abs=1e-300)
    assert vir[3] == pytest.approx(g[0][0] * dv[1], rel=1e-12, abs=1e-300)
    assert vir[2] == 0.0 and vir[4] == 0.0


def _cluster(seed=3, n=60):
    s = systems.cluster(n, seed=seed)
    rng = np.random.default_rng(seed)
    q = rng.uniform(-0.6, 0.6, s.n)
    q -= q.mean()
    return s, q


def test_gradient_is_the_energy_derivative(oracle_lib, ffield):
    s, q = _cluster()
    ev, ec, vir, g = _energy_grad(oracle_lib, ffield, s.elements, s.x, s.type, q)
    assert abs(ev) > 1 and abs(ec) > 1
    h = 1e-5
    for i, k in ((0, 0), (7, 1), (23, 2), (41, 0)):
        xp, xm = s.x.copy(), s.x.copy()
        xp[i, k] += h
        xm[i, k] -= h
        ep = sum(_energy_grad(oracle_lib, ffield, s.elements, xp, s.type, q)[:2])
        em = sum(_energy_grad(oracle_lib, ffield, s.elements, xm, s.type, q)[:2])
        fd = (ep - em) / (2 * h)
        assert fd == pytest.approx(g[i, k], rel=1e-7, abs=1e-7), (i, k)


def test_gradients_sum_to_zero(oracle_lib, ffield):
    s, q = _cluster(seed=5, n=200)
    _, _, _, g = _energy_grad(oracle_lib, ffield, s.elements, s.x, s.type, q)
    assert np.abs(g.sum(0)).max() <= 1e-10 * np.abs(g).max()


def inner_wall_ffield(ffield, path):
    """The force field with inner-wall parameters on every atom (vdw_type 3):
    rcore2, ecore2, acore2 are the last three numbers of each atom's 4th line."""
    lines = open(ffield).read().splitlines()
    k = 1
    ngp = int(lines[k].split()[0])
    k += 1 + ngp
    nat = int(lines[k].split()[0])
    k += 4
    for a in range(nat):
        t = lines[k + 3].split()
        t[5], t[6], t[7] = "1.6000", "0.0300", "13.5000"
        lines[k + 3] = " " + " ".join(f"{float(v):9.4f}" for v in t)
        k += 4
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return str(path)


def test_inner_wall_closed_form_and_gradient(oracle_lib, ffield, tmp_path):
    ffw = inner_wall_ffield(ffield, tmp_path / "ffield_wall.reax")
    ff = FField(ffw)
    assert ff.vdw_type() == 3
    o = oracle_lib.Oracle(ffw, ["H", "O"], threads=1)
    assert o.scalar("vdw_type") == 3
    els = ["C", "H", "O", "N"]
    for a, b, r in (("H", "O", 1.2), ("O", "O", 2.6), ("C", "N", 4.0)):
        x = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, r]])
        ev, ec, _, g = _energy_grad(oracle_lib, ffw, els, x, [els.index(a) + 1, els.index(b) + 1], [0.2, -0.1])
        Ev, Ec = ff.nonbonded_pair_energy(a, b, r, 0.2, -0.1)
        assert ev == pytest.approx(Ev, rel=1e-12) and ec == pytest.approx(Ec, rel=1e-12)
        h = 1e-6
        dEdr = (sum(ff.nonbonded_pair_energy(a, b, r + h, 0.2, -0.1)) -
                sum(ff.nonbonded_pair_energy(a, b, r - h, 0.2, -0.1))) / (2 * h)
        assert g[0, 2] == pytest.approx(-dEdr, rel=1e-6)


def test_missing_vdw_method_is_an_error(oracle_lib, ffield, tmp_path, gpu_lib):
    """An element with neither shielding (gamma_w > 0.5) nor an inner wall is
    rejected (reaxc_ffield_sunway.cpp:286-291), by the oracle and the library."""
    lines = open(ffield).read().splitlines()
    k = 1 + 1 + int(lines[1].split()[0])
    k += 4                                      # first atom's line 2 holds gamma_w at position 1
    t = lines[k + 1].split()
    t[1] = "0.1000"
    lines[k + 1] = " " + " ".join(f"{float(v):9.4f}" for v in t)
    bad = tmp_path / "ffield_bad.reax"
    bad.write_text("\n".join(lines) + "\n")
    with pytest.raises(RuntimeError, match="shielding or inner-wall"):
        oracle_lib.Oracle(str(bad), ["H", "O"], threads=1)
    with pytest.raises(gpu_lib.SwrError, match="shielding or inner-wall"):
        gpu_lib.System(str(bad), ["H", "O"])


# ------------------------------------------------------------------ GPU parity
def _gpu_nonbonded(lmp):
    from sw_reaxff_amd import _lib
    h = lmp.system
    _lib.check(h.L.swr_vdw_coulomb_energy(h.h), "swr_vdw_coulomb_energy")
    h.sync()
    e = h.get("nonbonded_energy")
    return e[0], e[1], e[2:], h.get("fCdDelta").reshape(-1, 4)


def _compare(lmp, oracle_lib, ff_path, s, xa, owner):
    ev, ec, vir, fcd = _gpu_nonbonded(lmp)
    h = lmp.system
    qN = h.get("q")
    N = xa.shape[0]
    o = oracle_lib.Oracle(ff_path, s.elements, threads=8)
    o.set_atoms(s.n, xa, s.type[owner].astype(np.int32), s.tag[owner].astype(np.int32), owner.astype(np.int32),
                q=qN[:N])
    o.build_verlet(12.0)
    for st in ("write_reax_atoms", "write_reax_lists", "estimate_storages", "reset"):
        o.stage(st)
    oev, oec, ovir = o.vdw_coulomb()
    og = o.get("fCdDelta").reshape(-1, 4)[: s.n, :3]
    print(f"[nonbonded] n={s.n} E_vdW gpu {ev:.12e} oracle {oev:.12e}  E_ele gpu {ec:.12e} oracle {oec:.12e}",
          flush=True)
    assert ev == pytest.approx(oev, rel=1e-10)
    assert ec == pytest.approx(oec, rel=1e-10)
    assert np.max(np.abs(vir - ovir)) <= 1e-10 * np.max(np.abs(ovir))
    ea = lmp.system.get("nonbonded_eatom")[: s.n]
    assert ea.sum() == pytest.approx(ev + ec, rel=1e-12)
    g = fcd[: s.n, :3]
    assert np.max(np.abs(g - og)) <= 1e-10 * np.max(np.abs(og))


@pytest.mark.gpu
def test_gpu_nonbonded_water_box(gpu_lib, oracle_lib, ffield):
    from sw_reaxff_amd.lammps import Lammps, standard_script
    s = systems.water_box(648, seed=3)
    lmp = Lammps()
    lmp.commands(standard_script(s.elements, ffield=ffield, tol=1e-10))
    lmp.create_atoms(s)
    lmp.setup()
    h = lmp.system
    xa = h.get("x").reshape(-1, 3)
    owner = h.get("owner")
    _compare(lmp, oracle_lib, ffield, s, xa, owner)
    # a second call adds the same gradient again (fCdDelta accumulates until Reset)
    before = h.get("fCdDelta").reshape(-1, 4)[: s.n, :3].copy()
    from sw_reaxff_amd import _lib
    _lib.check(h.L.swr_vdw_coulomb_energy(h.h), "swr_vdw_coulomb_energy")
    h.sync()
    after = h.get("fCdDelta").reshape(-1, 4)[: s.n, :3]
    assert np.allclose(after, 2 * before, rtol=1e-12, atol=1e-12)
    lmp.run(2)                                   # Reset zeroes it; the pass runs on the new lists
    _compare(lmp, oracle_lib, ffield, s, h.get("x").reshape(-1, 3), h.get("owner"))


@pytest.mark.gpu
def test_gpu_nonbonded_inner_wall_cluster(gpu_lib, oracle_lib, ffield, tmp_path):
    from sw_reaxff_amd.lammps import Lammps, standard_script
    ffw = inner_wall_ffield(ffield, tmp_path / "ffield_wall.reax")
    s = systems.cluster(700, seed=9)
    lmp = Lammps()
    lmp.commands(standard_script(s.elements, ffield=ffw, tol=1e-10))
    lmp.create_atoms(s)
    lmp.setup()
    h = lmp.system
    xa = h.get("x").reshape(-1, 3)
    _compare(lmp, oracle_lib, ffw, s, xa, h.get("owner"))
    fcd = h.get("fCdDelta").reshape(-1, 4)[: s.n, :3]
    assert np.abs(fcd.sum(0)).max() <= 1e-9 * np.abs(fcd).max()
