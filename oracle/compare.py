"""Whole-state comparison of the HIP path with the oracle (test
infrastructure: the checker).  Only the test suite and the benchmark's
CPU-baseline leg, which runs the oracle on the GPU's own state, use it.

Vectorised over all atoms, so the same checks run on a 648-water box and on
a 96,000-atom box.  The oracle is fed the GPU's atoms, Verlet list (so both
sides sum in one order), capacities and the QEq history the step started
from; `h` is any object with the get(name) / scalar(name) mirrors of the
library's system handle (swr_get / swr_get_scalar).
"""
from __future__ import annotations

import numpy as np

from . import pyoracle

# lists bit-exact as sets, bond orders and their coefficients 1e-10 relative
BO_RTOL = 1e-10


def rel_err(a, b, floor=1e-300):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.maximum(np.abs(b), floor)
    return np.max(np.abs(a - b) / den) if a.size else 0.0


def rel_err_scaled(a, b, frac=1e-5, unit=0.0):
    """rel_err with a floor at frac times the larger of the array's largest
    magnitude and unit: entries that are differences of much larger terms
    (Delta' = sum BO - valency near zero, a coefficient that cancels) are
    compared absolutely at rtol * frac of that magnitude, a few units in its
    last place.  unit = 1 for per-atom sums of bond orders and the valences
    and lone-pair counts they are compared with, which are of order one
    whatever the size of their difference."""
    b = np.asarray(b, dtype=np.float64)
    scale = max(float(np.abs(b).max()) if b.size else 0.0, unit)
    return rel_err(a, b, max(frac * scale, 1e-300))


def worst_entry(a, b, floor=1e-300):
    """The entry that sets rel_err: (index, a, b, largest |b|)."""
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    if not a.size:
        return None
    k = int(np.argmax(np.abs(a - b) / np.maximum(np.abs(b), floor)))
    return (k, float(a[k]), float(b[k]), float(np.abs(b).max()))


def atoms_of(h):
    x = h.get("x").reshape(-1, 3)
    return dict(n=h.scalar("n"), N=h.scalar("N"), x=x, type=h.get("type"), tag=h.get("tag"),
                owner=h.get("owner"), q=h.get("q"))


def verlet_of(h):
    off = h.get("verlet_offsets")
    lst = h.get("verlet_list")
    return off, lst[: off[-1]]


BO_FIELDS = ("BO_s", "dBOp", "dln_BOp_s", "dln_BOp_pi", "dln_BOp_pi2", "C1dbo", "C2dbo", "C3dbo", "C1dbopi",
             "C2dbopi", "C3dbopi", "C4dbopi", "C1dbopi2", "C2dbopi2", "C3dbopi2", "C4dbopi2")
WS_ARRAYS = ("bo_dboc", "dDeltap_self", "Deltap", "Deltap_boc", "Delta", "Delta_e", "Delta_val", "vlpex", "nlp",
             "nlp_temp", "Clp", "Delta_lp", "dDelta_lp", "Delta_lp_temp", "dDelta_lp_temp")


def ranges(idx, end):
    """Concatenated slot indices of per-atom ranges [idx[i], end[i])."""
    idx = np.asarray(idx, np.int64)
    cnt = np.asarray(end, np.int64) - idx
    first = np.concatenate(([0], np.cumsum(cnt)[:-1])) if cnt.size else cnt
    return np.repeat(idx - first, cnt) + np.arange(int(cnt.sum()), dtype=np.int64)


def feed_oracle(o, h, hist=None):
    """Gives the oracle the GPU's state of the step just run: atoms, Verlet
    list, the capacities its Validate_Lists wrote (each >= the step's counts)
    and, when given, the QEq history the step started from."""
    A = atoms_of(h)
    o.set_atoms(A["n"], A["x"], A["type"], A["tag"], A["owner"])
    o.set_verlet(*verlet_of(h))
    o.set("num_bonds", np.ascontiguousarray(h.get("num_bonds")[: A["N"]]))
    o.set("num_hbonds", np.ascontiguousarray(h.get("num_hbonds")[: A["N"]]))
    if hist is not None:
        o.set("s_hist", np.ascontiguousarray(hist[0]))
        o.set("t_hist", np.ascontiguousarray(hist[1]))
    return A


def history_of(h):
    return h.get("s_hist").copy(), h.get("t_hist").copy()


def compare_state(h, o, q_atol, bo_rtol=BO_RTOL, far=True):
    """GPU state after a step against the oracle's step on the same inputs:
    far, bond and hbond lists equal per atom (bitwise where both compute the
    same expression in the same order), bond orders and their coefficients
    within bo_rtol relative, charges within q_atol.  Returns the measured
    differences."""
    st = {}
    n, N = h.scalar("n"), h.scalar("N")
    gq, oq = h.get("q")[:N], o.get("q")[:N]
    st["dq"] = float(np.max(np.abs(gq - oq)))
    assert st["dq"] <= q_atol, st
    gfi, gfe = h.get("far_index")[:N], h.get("far_end")[:N]
    ofi, ofe = o.get("far_index")[:N], o.get("far_end")[:N]
    assert np.array_equal(gfe - gfi, ofe - ofi), "far-list counts differ"
    st["far"] = int((gfe - gfi).astype(np.int64).sum())
    if far:
        gf, of = h.get("far")[ranges(gfi, gfe)], o.get("far")[ranges(ofi, ofe)]
        for f in ("nbr", "type", "orig_id", "d", "dvec"):
            assert np.array_equal(gf[f], of[f]), f"far field {f}"
        assert not np.any(gf["rel_box"])
        st["far_dq"] = float(np.max(np.abs(gf["q"] - of["q"]), initial=0.0))
        assert st["far_dq"] <= q_atol
        del gf, of
    gb, ge = h.get("bond_index")[:N], h.get("bond_end")[:N]
    ob, oe = o.get("bond_index")[:N], o.get("bond_end")[:N]
    assert np.array_equal(ge - gb, oe - ob), "bond counts differ"
    gs, os_ = ranges(gb, ge), ranges(ob, oe)
    st["bonds"] = int(gs.size)
    gB, oB = h.get("bonds")[gs], o.get("bonds")[os_]
    for f in ("nbr", "d", "dvec"):
        assert np.array_equal(gB[f], oB[f]), f"bond field {f}"
    own = np.repeat(np.arange(N), (ge - gb).astype(np.int64))
    assert np.array_equal(gB["dbond_index"] - gb[own], oB["dbond_index"] - ob[own])
    assert np.array_equal(gB["sym_index"] - gb[gB["nbr"]], oB["sym_index"] - ob[oB["nbr"]])
    # bond orders strictly relative (floor 1e-12: BO'-bo_cut of a weak bond);
    # their coefficients relative down to 1e-5 of each array's largest
    # magnitude and the per-atom sums down to 1e-5 of the larger of that and
    # one, absolute below it
    worst = {"BO": rel_err(h.get("BO")[gs], o.get("BO")[os_], 1e-12),
             "BOpi": rel_err(h.get("BOpi").reshape(-1, 2)[gs], o.get("BOpi").reshape(-1, 2)[os_], 1e-12)}
    gbo, obo = h.get("bo_data")[gs], o.get("bo_data")[os_]
    for f in BO_FIELDS:
        worst[f] = rel_err_scaled(gbo[f], obo[f])
    for nm in WS_ARRAYS:
        worst[nm] = rel_err_scaled(h.get(nm), o.get(nm), unit=1.0)
    st["bo_rel"] = float(max(worst.values()))
    if st["bo_rel"] > bo_rtol:
        arrays = {"BO": (h.get("BO")[gs], o.get("BO")[os_])}
        arrays.update({f: (gbo[f], obo[f]) for f in BO_FIELDS})
        arrays.update({nm: (h.get(nm), o.get(nm)) for nm in WS_ARRAYS})
        bad = {k: (v, worst_entry(*arrays[k]) if k in arrays else None)
               for k, v in worst.items() if v > bo_rtol}
        raise AssertionError(f"relative differences beyond {bo_rtol}: {bad}")
    assert np.array_equal(h.get("bond_mark")[:N], o.get("bond_mark")[:N]), "bond_mark"
    assert np.array_equal(h.get("num_bonds")[:N], o.get("num_bonds")[:N]), "num_bonds"
    assert np.array_equal(h.get("num_hbonds")[:n], o.get("num_hbonds")[:n]), "num_hbonds"
    gH, oH = h.get("Hindex")[:n], o.get("Hindex")[:n]
    assert np.array_equal(gH, oH), "Hindex"
    hat = np.empty(int((gH >= 0).sum()), np.int64)
    hat[gH[gH >= 0]] = np.nonzero(gH >= 0)[0]
    gi, gE, oi, oE = h.get("hbond_index"), h.get("hbond_end"), o.get("hbond_index"), o.get("hbond_end")
    assert np.array_equal(gE - gi, oE - oi), "hbond counts differ"
    hs, hos = ranges(gi, gE), ranges(oi, oE)
    st["hbonds"] = int(hs.size)
    assert np.array_equal(h.get("hbonds")["nbr"][hs], o.get("hbonds")["nbr"][hos])
    hown = hat[np.repeat(np.arange(hat.size), (gE - gi).astype(np.int64))]
    assert np.array_equal(h.get("hbond_far")[hs] - gfi[hown], o.get("hbond_far")[hos] - ofi[hown])
    return st


def oracle_energy_forces(o, owner, n):
    """The oracle's full ReaxFF energy terms and forces on the local atoms,
    after its step (the harness order: bond/atom, valence, torsion, hydrogen
    bonds, non-bonded, then the bond-order chain rule)."""
    e = {}
    e["bond"], e["lp"], e["ov"], e["un"], e["pol"] = (float(v) for v in o.atom_energy())
    e["ang"], e["pen"], e["coa"] = (float(v) for v in o.valence_angles())
    e["tor"], e["con"] = (float(v) for v in o.torsion_angles())
    e["hb"] = float(o.hydrogen_bonds())
    ev, ec, _ = o.vdw_coulomb()
    e["vdw"], e["coul"] = float(ev), float(ec)
    o.dbond_forces()
    return e, pyoracle.forces(o.get("fCdDelta"), owner, n)


def gpu_energy_forces(h):
    """The same terms as the GPU harness left them after a step with forces."""
    e = {}
    e["vdw"], e["coul"] = (float(v) for v in h.get("nonbonded_energy")[:2])
    e["bond"], e["lp"], e["ov"], e["un"], e["pol"] = (float(v) for v in h.get("bonded_energy"))
    e["hb"] = float(h.get("hbond_energy")[0])
    e["ang"], e["pen"], e["coa"] = (float(v) for v in h.get("valence_energy"))
    e["tor"], e["con"] = (float(v) for v in h.get("torsion_energy"))
    return e, h.get("f").reshape(-1, 3)


def compare_energy_forces(h, o, owner, n):
    """Relative differences of every energy term (floor 1 kcal/mol) and of
    the forces (relative to the largest force component)."""
    eg, fg = gpu_energy_forces(h)
    eo, fo = oracle_energy_forces(o, owner, n)
    de = max(abs(eg[k] - eo[k]) / max(abs(eo[k]), 1.0) for k in eo)
    df = float(np.max(np.abs(fg - fo)) / np.max(np.abs(fo)))
    return float(de), df
