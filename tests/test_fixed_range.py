"""Range of the fixed-point accumulators of the bonded kernels.

The hydrogen-bond, valence and torsion kernels add the terms of atoms other
than the centre into pairs of 64-bit integer words (a fine word in units of
2^-40, a coarse word in units of 2^-20), so their sums are exact and do not
depend on the order the adds arrive in.  A pair holds sums up to 2^43 and
terms up to ~3.8e12; beyond that the kernel raises a status flag and the
next swr_sync returns SWR_ERR_NUMERIC rather than a wrapped sum.

The test system makes those limits reachable: an acceptor O on the x axis
and six O-H donors on a cone of half-angle 40 degrees around it, each H
2.6 A from the acceptor and pointing straight at it.  The acceptor is
bonded to nothing (every H-O and O-O bond order it sees is below 1e-3), so
its hydrogen-bond force is the sum of exactly six terms, one per donor.
With the H...O...H angle at 180 degrees the angular derivative vanishes and
each term points along the H-acceptor direction: by symmetry the six terms
have the same x component, F_x / 6, and their y and z components are at most
tan(40 deg) times that.  Every energy and force of the hydrogen-bond term is
proportional to p_hb1 (the Dehb column of the force field's hydrogen-bond
lines), so scaling p_hb1 by s scales every term by s:

  * s with F_x(s) = 2^22 and 0.6 * 2^43: no flag; the GPU's forces equal the
    oracle's (FP64 sums) at 1e-10 relative, so large terms are split between
    the words without loss;
  * s with F_x(s) = 1.2 * 2^43 and every term below 0.4 * 2^43: the sixth add
    wraps the acceptor's coarse x word, which atomicAdd's old value detects;
  * s with every term beyond the per-term limit: refused term by term.
"""
import math
import os

import numpy as np
import pytest

from sw_reaxff_amd import systems

from .test_bonded import run_stages

ALPHA = math.radians(40.0)
R_HK, R_OH = 2.6, 0.97
NDONOR = 6
WRAP = 2.0 ** 43


def cone_system():
    x = [[0.0, 0.0, 0.0]]
    types = [1]
    for m in range(NDONOR):
        phi = 2.0 * math.pi * m / NDONOR + 0.3
        u = np.array([math.cos(ALPHA), math.sin(ALPHA) * math.cos(phi), math.sin(ALPHA) * math.sin(phi)])
        x += [list(R_HK * u), list((R_HK + R_OH) * u)]
        types += [2, 1]
    x = np.array(x) + 15.0
    return systems.System(x=x, v=np.zeros_like(x), type=np.array(types, np.int32),
                          tag=np.arange(1, len(types) + 1, dtype=np.int32), elements=["O", "H"],
                          lo=np.zeros(3), hi=np.full(3, 30.0), periodic=False)


def scaled_ffield(src, dst, s):
    """A copy of the force field with p_hb1 of every hydrogen-bond line times s."""
    lines = open(src).read().splitlines()
    k = next(i for i, l in enumerate(lines) if "hydrogen bonds" in l.lower())
    nhb = int(lines[k].split()[0])
    for i in range(k + 1, k + 1 + nhb):
        t = lines[i].split()
        t[4] = f"{float(t[4]) * s:.17g}"
        lines[i] = " ".join(t)
    with open(dst, "w") as f:
        f.write("\n".join(lines) + "\n")
    return dst


def _oracle_hb(oracle_lib, ffield, s):
    o = oracle_lib.Oracle(ffield, s.elements, threads=1)
    o.set_atoms(s.n, s.x, s.type.astype(np.int32), s.tag.astype(np.int32), np.arange(s.n, dtype=np.int32))
    o.build_verlet(12.0)
    run_stages(o)
    o.atom_energy()
    e = o.hydrogen_bonds()
    return o, e, o.get("fCdDelta").reshape(-1, 4)[: s.n, :3].copy()


def test_scaled_ffield_scales_the_hydrogen_bond_terms(oracle_lib, ffield, tmp_path):
    """The geometry's pins: the acceptor has no bond, its force has equal
    shares from the six donors, and scaling p_hb1 scales energy and forces."""
    s = cone_system()
    o, e1, g1 = _oracle_hb(oracle_lib, ffield, s)
    bi, be, BO = o.get("bond_index"), o.get("bond_end"), o.get("BO")
    assert np.all(BO[bi[0]:be[0]] < 1e-3)                 # the acceptor is bonded to nothing
    assert e1 < -1.0 and g1[0, 0] != 0.0
    # the six donors are images of each other under rotation about x
    gx = g1[2::2, 0]                                      # donor O atoms: x components equal
    assert np.ptp(gx) <= 1e-12 * np.abs(gx).max()
    assert np.hypot(g1[0, 1], g1[0, 2]) <= 1e-10 * abs(g1[0, 0])   # the acceptor's y, z cancel
    _, e7, g7 = _oracle_hb(oracle_lib, scaled_ffield(ffield, str(tmp_path / "ff7"), 7.0), s)
    assert e7 == pytest.approx(7.0 * e1, rel=1e-12)
    assert np.abs(g7 - 7.0 * g1).max() <= 1e-12 * np.abs(g7).max()


def _scale_for(oracle_lib, ffield, target_fx):
    _, _, g1 = _oracle_hb(oracle_lib, ffield, cone_system())
    return target_fx / abs(g1[0, 0])


def _gpu_hb(ffield_path, s):
    from sw_reaxff_amd import _lib
    from .parity import harness
    lmp = harness(s, ffield_path, tol=1e-10)
    h = lmp.system
    _lib.check(h.L.swr_atom_energy(h.h), "swr_atom_energy")
    _lib.check(h.L.swr_hydrogen_bonds(h.h), "swr_hydrogen_bonds")
    h.sync()
    return h.get("hbond_energy")[0], h.get("fCdDelta").reshape(-1, 4)[: s.n, :3].copy()


@pytest.mark.gpu
@pytest.mark.parametrize("target", [2.0 ** 22, 0.6 * WRAP])
def test_gpu_large_hydrogen_bond_terms_match_oracle(gpu_lib, oracle_lib, ffield, tmp_path, target):
    s = cone_system()
    sc = _scale_for(oracle_lib, ffield, target)
    ff = scaled_ffield(ffield, str(tmp_path / "ff"), sc)
    print(f"[fixed-range] target F_x {target:.3e} p_hb1 x {sc:.4e}", flush=True)
    e_g, g_g = _gpu_hb(ff, s)
    _, e_o, g_o = _oracle_hb(oracle_lib, ff, s)
    assert abs(g_o[0, 0]) == pytest.approx(target, rel=1e-9)
    assert e_g == pytest.approx(e_o, rel=1e-10)
    assert np.abs(g_g - g_o).max() <= 1e-10 * np.abs(g_o).max()


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["sum_wraps", "term_beyond_range"])
def test_gpu_accumulator_out_of_range_is_reported(gpu_lib, oracle_lib, ffield, tmp_path, case):
    from sw_reaxff_amd import _lib
    s = cone_system()
    if case == "sum_wraps":
        target = 1.2 * WRAP
        term = target / NDONOR / math.cos(ALPHA)          # each term's length
        assert term < 0.4 * WRAP                          # every add is in range; only their sum is not
    else:
        target = 8.0 * WRAP                               # each term ~2^43 / cos(40 deg), beyond ~3.8e12
    sc = _scale_for(oracle_lib, ffield, target)
    ff = scaled_ffield(ffield, str(tmp_path / "ff"), sc)
    print(f"[fixed-range] {case}: target F_x {target:.3e} p_hb1 x {sc:.4e}", flush=True)
    with pytest.raises(_lib.SwrError) as ei:
        _gpu_hb(ff, s)
    assert ei.value.code == 8, str(ei.value)              # SWR_ERR_NUMERIC
