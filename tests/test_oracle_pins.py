"""Known-answer and invariant tests that pin the FP64 oracle (CPU only).

The reference ships no golden vectors for this path, so the oracle is pinned
by what the mathematics fixes, each chosen so that a plausible slip (a dropped
term, a wrong sign or index, a transposed parameter, a mis-read force-field
column) fails one of them:

  a1   brute-force O(N^2) pair enumeration == the cell list (exact sets),
       and local pair symmetry;
  a3   far entries: d == sqrt(dx^2+dy^2+dz^2) bit for bit, dvec == x_j - x_i,
       the set == Verlet pairs within nonb_cut;
  a5   BO' of isolated 2-atom pairs from the raw force-field text
       (tests/ffield_kat.py), finite-difference dBOp, bond set membership;
  a6   hbond set {local H, O/N within hbond_cut}, ptr -> far entry of j;
  a9   sym_index involution, BO_ij == BO_ji, sum_j BO_ij == Delta_i + valency_i,
       2-atom f4 f5 correction in closed form;
  a10  element-wise closed forms of the lone-pair terms;
  a11/a13  taper identities and the closed-form 7th-order polynomial;
  a14  SpMV == dense matvec;
  a15/a16  2-atom electronegativity equalisation in closed form, a numpy
       dense solve of the oracle's own H, sum q == 0.
"""
import math

import numpy as np
import pytest

from sw_reaxff_amd import systems

from .ffield_kat import FField

BOND_CUT, HBOND_CUT, NONB_CUT, SWB = 5.0, 7.5, 10.0, 10.0


@pytest.fixture(scope="module")
def ff(ffield):
    return FField(ffield)


def ieee_d(x, i=0, j=1):
    """d as the path computes it: sqrt(dx*dx + dy*dy + dz*dz), no FMA."""
    dv = x[j] - x[i]
    return float(np.sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]))


def vacuum(oracle_lib, ffield, elements, x, types, tol=1e-10, step=True):
    """An oracle system of isolated atoms (no periodic images)."""
    x = np.asarray(x, dtype=np.float64)
    n = x.shape[0]
    o = oracle_lib.Oracle(ffield, elements, tol=tol, threads=1)
    o.set_atoms(n, x, np.asarray(types, np.int32), np.arange(1, n + 1, dtype=np.int32),
                np.arange(n, dtype=np.int32))
    o.build_verlet(12.0)
    if step:
        o.step()
    return o


def periodic_c0(oracle_lib, ffield, seed=3, nmol=648, tol=1e-6, threads=8):
    s = systems.water_box(nmol, seed=seed)
    xa, owner = systems.borders(s.x, s.lo, s.hi, 12.0)
    N = xa.shape[0]
    types = s.type[owner]
    tags = s.tag[owner]
    o = oracle_lib.Oracle(ffield, s.elements, tol=tol, threads=threads)
    o.set_atoms(s.n, xa, types, tags, owner)
    return s, o, xa, owner, N


# ------------------------------------------------------------------ params
def test_force_field_reader_matches_raw_text(oracle_lib, ffield, ff):
    o = oracle_lib.Oracle(ffield, ["C", "H", "O", "N"], threads=1)
    assert o.control("bo_cut") == pytest.approx(ff.bo_cut, rel=0, abs=1e-18)
    assert ff.bo_cut == pytest.approx(1e-4)
    for e, el in enumerate(["C", "H", "O", "N"]):
        sb = o.sbp(e)
        A = ff.atom(el)
        for k in ("r_s", "valency", "mass", "gamma", "r_pi", "valency_e", "valency_boc", "chi",
                  "p_hbond", "r_pi_pi", "b_o_131", "b_o_132", "b_o_133", "valency_val"):
            assert sb[k] == A[k], (el, k)
        assert sb["eta"] == ff.eta(el), el         # eta = 2 x the file value
        for f, el2 in enumerate(["C", "H", "O", "N"]):
            tb, P = o.tbp(e, f), ff.pair(el, el2)
            for k in ("p_bo1", "p_bo2", "p_bo3", "p_bo4", "p_bo5", "p_bo6", "r_s", "r_p", "r_pp", "v13cor",
                      "ovc", "p_boc3", "p_boc4", "p_boc5"):
                assert tb[k] == pytest.approx(P[k], rel=1e-15, abs=0), (el, el2, k)


# ------------------------------------------------------------------ taper (a11, a13)
def test_taper_identities(oracle_lib, ffield, ff):
    o = oracle_lib.Oracle(ffield, ["H", "O"], threads=1)
    T = o.taper()
    p = np.polynomial.Polynomial(T)
    assert p(0.0) == pytest.approx(1.0, abs=1e-15)
    assert p(SWB) == pytest.approx(0.0, abs=1e-12)
    for k in (1, 2, 3):
        assert p.deriv(k)(SWB) == pytest.approx(0.0, abs=1e-10), k
        assert p.deriv(k)(0.0) == pytest.approx(0.0, abs=1e-12), k
    for r in np.linspace(0, SWB, 17):
        assert p(r) == pytest.approx(ff.taper_closed_form(r), abs=1e-13)


# ------------------------------------------------------------------ neighbour list (a1)
def test_cell_list_equals_brute_force(oracle_lib, ffield):
    _, o, xa, owner, N = periodic_c0(oracle_lib, ffield, nmol=256)
    o.build_verlet(12.0, brute=True)
    ob, lb = o.get("verlet_offsets"), o.get("verlet_list")
    o.build_verlet(12.0, brute=False)
    oc, lc = o.get("verlet_offsets"), o.get("verlet_list")
    assert ob[-1] == oc[-1] > 0
    for i in range(N):
        a = np.sort(lb[ob[i]:ob[i + 1]])
        b = np.sort(lc[oc[i]:oc[i + 1]])
        assert np.array_equal(a, b), i
        assert i not in set(b.tolist())
    # distances: every listed pair is within the cut, every unlisted one is not
    i = 17
    d2 = np.sum((xa - xa[i]) ** 2, axis=1)
    want = set(np.nonzero(d2 <= 144.0)[0].tolist()) - {i}
    assert set(lc[oc[i]:oc[i + 1]].tolist()) == want


def test_local_pairs_are_symmetric(oracle_lib, ffield):
    s, o, xa, owner, N = periodic_c0(oracle_lib, ffield, nmol=256)
    o.build_verlet(12.0)
    off, lst = o.get("verlet_offsets"), o.get("verlet_list")
    n = s.n
    pairs = set()
    for i in range(n):
        for j in lst[off[i]:off[i + 1]]:
            if j < n:
                pairs.add((i, int(j)))
    assert all((j, i) in pairs for (i, j) in pairs)


# ------------------------------------------------------------------ far list (a3)
def test_far_entries_exact(oracle_lib, ffield):
    s, o, xa, owner, N = periodic_c0(oracle_lib, ffield, nmol=128)
    o.build_verlet(12.0)
    o.step()
    far, fi, fe = o.get("far"), o.get("far_index"), o.get("far_end")
    off, lst = o.get("verlet_offsets"), o.get("verlet_list")
    q = o.get("q")
    ffidx = np.array([{"C": 0, "H": 1, "O": 2, "N": 3}[e] for e in s.elements])
    pk = o.get("packed_atoms")
    assert np.array_equal(pk["type"], ffidx[s.type[owner] - 1])
    assert np.array_equal(pk["x"], xa) and np.array_equal(pk["orig_id"], s.tag[owner])
    for i in range(0, N, 7):
        e = far[fi[i]:fe[i]]
        j = e["nbr"]
        dv = xa[j] - xa[i]
        assert np.array_equal(e["dvec"], dv)
        d = np.sqrt(dv[:, 0] * dv[:, 0] + dv[:, 1] * dv[:, 1] + dv[:, 2] * dv[:, 2])
        assert np.array_equal(e["d"], d)
        assert np.array_equal(e["orig_id"], s.tag[owner[j]])
        assert np.array_equal(e["type"], ffidx[s.type[owner[j]] - 1])
        assert np.all(e["rel_box"] == 0)
        assert np.array_equal(e["q"], q[j])
        vj = lst[off[i]:off[i + 1]]
        dd = np.sum((xa[vj] - xa[i]) ** 2, axis=1)
        assert set(j.tolist()) == set(vj[dd <= NONB_CUT ** 2].tolist())


# ------------------------------------------------------------------ BO' (a5)
@pytest.mark.parametrize("a,b,d", [("H", "O", 0.98), ("H", "O", 1.6), ("H", "H", 0.76), ("O", "O", 1.25),
                                   ("C", "O", 1.22), ("C", "C", 1.40), ("C", "N", 1.35), ("N", "O", 1.21)])
def test_uncorrected_bond_order_two_atoms(oracle_lib, ffield, ff, a, b, d):
    els = [a] if a == b else [a, b]
    types = [1, 1] if a == b else [1, 2]
    x = np.array([[20.0, 20.0, 20.0], [20.0 + d * 0.6, 20.0 + d * 0.8, 20.0]])
    d = ieee_d(x)
    o = vacuum(oracle_lib, ffield, els, x, types, step=False)
    o.stage("write_reax_atoms")
    o.stage("write_reax_lists")
    o.stage("estimate_storages")
    o.stage("reset")
    o.stage("write_reax_lists")
    assert o.stage("init_forces") == 0
    bonds, bi, be = o.get("bonds"), o.get("bond_index"), o.get("bond_end")
    s, pi, pp = ff.bop(a, b, d)
    bop = s + pi + pp
    if bop < ff.bo_cut:
        assert be[0] == bi[0] and be[1] == bi[1]
        return
    assert be[0] - bi[0] == 1 and be[1] - bi[1] == 1
    k = bi[0]
    assert bonds["nbr"][k] == 1
    BO, BOpi, bo = o.get("BO"), o.get("BOpi").reshape(-1, 2), o.get("bo_data")
    assert BO[k] == pytest.approx(bop - ff.bo_cut, rel=1e-13)
    assert bo["BO_s"][k] == pytest.approx(s - ff.bo_cut, rel=1e-13, abs=1e-16)
    assert BOpi[k, 0] == pytest.approx(pi, rel=1e-13, abs=1e-300)
    assert BOpi[k, 1] == pytest.approx(pp, rel=1e-13, abs=1e-300)
    # dBOp = d BO'/d x_i (dvec = x_j - x_i): central differences on the text formula
    h = 1e-6
    u = (x[1] - x[0]) / d
    fd = -(sum(ff.bop(a, b, d + h)) - sum(ff.bop(a, b, d - h))) / (2 * h) * u
    assert np.allclose(bo["dBOp"][k], fd, rtol=1e-6, atol=1e-9)
    # dDeltap_self[i] accumulates dBOp; bo_dboc sums go to Deltap + valency
    assert np.allclose(o.get("dDeltap_self").reshape(-1, 3)[0], bo["dBOp"][k], rtol=1e-14, atol=0)


def _fd_corrected_two_atom(ff, a, b, d):
    """Corrected BO of an isolated pair (f1 = 1 unless ovc; f4 f5 with v13cor),
    from the published ReaxFF bond-order correction."""
    s, pi, pp = ff.bop(a, b, d)
    cut = ff.bo_cut
    BO = s + pi + pp - cut
    BOs, BOpi, BOpp = s - cut, pi, pp
    A, B, t = ff.atom(a), ff.atom(b), ff.pair(a, b)
    if t["ovc"] < 0.001 and t["v13cor"] < 0.001:
        return BO, BOpi, BOpp
    Dp_i, Dp_j = BO - A["valency"], BO - B["valency"]
    Dboc_i, Dboc_j = BO - A["valency_val"], BO - B["valency_val"]
    if t["ovc"] >= 0.001:
        p1, p2 = ff.gp[0], ff.gp[1]
        exp_p1i, exp_p2i = math.exp(-p1 * Dp_i), math.exp(-p2 * Dp_i)
        exp_p1j, exp_p2j = math.exp(-p1 * Dp_j), math.exp(-p2 * Dp_j)
        f2 = exp_p1i + exp_p1j
        f3 = -1.0 / p2 * math.log(0.5 * (exp_p2i + exp_p2j))
        f1 = 0.5 * ((A["valency"] + f2) / (A["valency"] + f2 + f3) + (B["valency"] + f2) / (B["valency"] + f2 + f3))
    else:
        f1 = 1.0
    if t["v13cor"] >= 0.001:
        f4 = 1.0 / (1.0 + math.exp(-t["p_boc3"] * (t["p_boc4"] * BO * BO - Dboc_i) + t["p_boc5"]))
        f5 = 1.0 / (1.0 + math.exp(-t["p_boc3"] * (t["p_boc4"] * BO * BO - Dboc_j) + t["p_boc5"]))
    else:
        f4 = f5 = 1.0
    A0 = f1 * f4 * f5
    out = [BO * A0, BOpi * A0 * f1, BOpp * A0 * f1]
    # components below 1e-10 are flushed to zero (reaxc_bond_orders_sw64.c:341-355)
    return tuple(v if v >= 1e-10 else 0.0 for v in out)


@pytest.mark.parametrize("a,b,d", [("H", "O", 0.98), ("O", "O", 1.21), ("C", "C", 1.40), ("H", "H", 0.9),
                                   ("C", "H", 1.09)])
def test_corrected_bond_order_two_atoms(oracle_lib, ffield, ff, a, b, d):
    els = [a] if a == b else [a, b]
    types = [1, 1] if a == b else [1, 2]
    x = np.array([[20.0, 20.0, 20.0], [20.0, 20.0 + d, 20.0]])
    d = ieee_d(x)
    o = vacuum(oracle_lib, ffield, els, x, types)
    BO, BOpi = o.get("BO"), o.get("BOpi").reshape(-1, 2)
    bi = o.get("bond_index")
    want = _fd_corrected_two_atom(ff, a, b, d)
    for i in (0, 1):
        k = bi[i]
        assert BO[k] == pytest.approx(want[0], rel=1e-12), (i, "BO")
        assert BOpi[k, 0] == pytest.approx(want[1], rel=1e-12, abs=1e-14)
        assert BOpi[k, 1] == pytest.approx(want[2], rel=1e-12, abs=1e-14)
    # Section 3: Delta_i = sum_j BO_ij - valency_i
    D = o.get("Delta")
    assert D[0] == pytest.approx(want[0] - ff.atom(a)["valency"], rel=1e-12, abs=1e-14)


# ------------------------------------------------------------------ bulk invariants (a5-a10)
@pytest.fixture(scope="module")
def c0_oracle(oracle_lib, ffield):
    s, o, xa, owner, N = periodic_c0(oracle_lib, ffield, nmol=256, tol=1e-10)
    o.build_verlet(12.0)
    o.step()
    return s, o, xa, owner, N


def test_bond_set_and_order(c0_oracle, ff):
    s, o, xa, owner, N = c0_oracle
    bonds, bi, be = o.get("bonds"), o.get("bond_index"), o.get("bond_end")
    far, fi, fe = o.get("far"), o.get("far_index"), o.get("far_end")
    els = s.elements
    for i in range(0, N, 5):
        nb = bonds["nbr"][bi[i]:be[i]]
        assert np.all(np.diff(nb) > 0), i                 # sorted by nbr, no repeats
        e = far[fi[i]:fe[i]]
        ti = els[s.type[owner[i]] - 1]
        want = [int(e["nbr"][k]) for k in range(len(e)) if e["d"][k] <= BOND_CUT and
                sum(ff.bop(ti, els[s.type[owner[e["nbr"][k]]] - 1], e["d"][k])) >= ff.bo_cut]
        assert sorted(want) == nb.tolist(), i


def test_sym_index_and_bond_order_symmetry(c0_oracle):
    s, o, xa, owner, N = c0_oracle
    bonds, bi, be = o.get("bonds"), o.get("bond_index"), o.get("bond_end")
    BO = o.get("BO")
    sym = bonds["sym_index"]
    owner_of = np.repeat(np.arange(N), be - bi)
    sel = np.concatenate([np.arange(bi[i], be[i]) for i in range(N)])
    assert np.array_equal(sym[sym[sel]], sel)
    assert np.array_equal(bonds["nbr"][sym[sel]], owner_of)
    assert np.allclose(BO[sel], BO[sym[sel]], rtol=1e-13, atol=0)


def test_delta_is_bond_order_sum(c0_oracle, ff):
    s, o, xa, owner, N = c0_oracle
    bi, be, BO = o.get("bond_index"), o.get("bond_end"), o.get("BO")
    D, De = o.get("Delta"), o.get("Delta_e")
    val = np.array([ff.atom(e)["valency"] for e in s.elements])[s.type[owner] - 1]
    vale = np.array([ff.atom(e)["valency_e"] for e in s.elements])[s.type[owner] - 1]
    tot = np.array([BO[bi[i]:be[i]].sum() for i in range(N)])
    assert np.allclose(D[:N], tot - val, rtol=0, atol=1e-12)
    assert np.allclose(De[:N], tot - vale, rtol=0, atol=1e-12)


def test_lone_pair_terms_closed_form(c0_oracle, ff):
    s, o, xa, owner, N = c0_oracle
    De = o.get("Delta_e")[:N]
    p_lp1 = ff.gp[15]
    vlpex = De - 2.0 * np.trunc(De / 2.0)
    explp1 = np.exp(-p_lp1 * (2.0 + vlpex) ** 2)
    nlp = explp1 - np.trunc(De / 2.0)
    nlp_opt = np.array([0.5 * (ff.atom(e)["valency_e"] - ff.atom(e)["valency"]) for e in s.elements])[
        s.type[owner] - 1]
    assert np.allclose(o.get("vlpex")[:N], vlpex, rtol=0, atol=1e-14)
    assert np.allclose(o.get("nlp")[:N], nlp, rtol=0, atol=1e-14)
    assert np.allclose(o.get("Delta_lp")[:N], nlp_opt - nlp, rtol=0, atol=1e-14)
    Clp = 2.0 * p_lp1 * explp1 * (2.0 + vlpex)
    assert np.allclose(o.get("Clp")[:N], Clp, rtol=1e-14, atol=1e-300)
    heavy = np.array([ff.atom(e)["mass"] > 21.0 for e in s.elements])[s.type[owner] - 1]
    assert not heavy.any()          # water: nlp_temp = nlp, dDelta_lp_temp = Clp
    assert np.allclose(o.get("nlp_temp")[:N], nlp, rtol=0, atol=1e-14)


def test_hbond_set_and_pointers(c0_oracle):
    s, o, xa, owner, N = c0_oracle
    far, fi, fe = o.get("far"), o.get("far_index"), o.get("far_end")
    hb, hi, he, Hindex = o.get("hbonds"), o.get("hbond_index"), o.get("hbond_end"), o.get("Hindex")
    hfar = o.get("hbond_far")
    ptype = {"H": 1, "O": 2}
    n = s.n
    el = np.array(s.elements)[s.type[owner] - 1]
    nH = 0
    for i in range(n):
        if el[i] != "H":
            continue
        nH += 1
        H = Hindex[i]
        got = hb["nbr"][hi[H]:he[H]]
        e = far[fi[i]:fe[i]]
        want = e["nbr"][(e["d"] <= HBOND_CUT) & np.array([ptype.get(el[j], 0) == 2 for j in e["nbr"]])]
        assert got.tolist() == want.tolist(), i           # far-list order
        slots = hfar[hi[H]:he[H]]
        assert np.array_equal(far["nbr"][slots], got)
        assert np.all((slots >= fi[i]) & (slots < fe[i]))
    assert nH == 2 * (n // 3)


# ------------------------------------------------------------------ QEq (a12-a16)
def test_two_atom_charge_equalisation(oracle_lib, ffield, ff):
    for (a, b, d) in [("H", "O", 1.0), ("C", "O", 2.5), ("H", "N", 6.0)]:
        x = np.array([[20.0, 20.0, 20.0], [20.0 + d, 20.0, 20.0]])
        d = ieee_d(x)
        o = vacuum(oracle_lib, ffield, [a, b], x, [1, 2])
        q = o.get("q")
        h = ff.h_offdiag(a, b, d)
        qa = (ff.atom(b)["chi"] - ff.atom(a)["chi"]) / (ff.eta(a) + ff.eta(b) - 2 * h)
        assert q[0] == pytest.approx(qa, rel=1e-9), (a, b, d)
        assert q[1] == pytest.approx(-qa, rel=1e-9), (a, b, d)
        Hv = o.get("H_val")
        assert Hv.size == 2 and np.allclose(Hv, h, rtol=1e-12)


def _dense_H(o, owner, n):
    f, nn, j, v = o.get("H_firstnbr"), o.get("H_numnbrs"), o.get("H_jlist"), o.get("H_val")
    Hd = np.zeros((n, n))
    for i in range(n):
        np.add.at(Hd[i], owner[j[f[i]:f[i] + nn[i]]], v[f[i]:f[i] + nn[i]])
    return Hd


def test_dense_solve_matches_cg(c0_oracle):
    s, o, xa, owner, N = c0_oracle
    n = s.n
    Hd = _dense_H(o, owner, n)
    assert np.allclose(Hd, Hd.T, rtol=1e-13, atol=1e-12)
    A = Hd + np.diag(1.0 / o.get("Hdia_inv")[:n])
    sd = np.linalg.solve(A, o.get("b_s")[:n])
    td = np.linalg.solve(A, o.get("b_t")[:n])
    assert np.max(np.abs(o.get("s")[:n] - sd)) <= 1e-8 * np.max(np.abs(sd))
    assert np.max(np.abs(o.get("t")[:n] - td)) <= 1e-8 * np.max(np.abs(td))
    qd = sd - sd.sum() / td.sum() * td
    q = o.get("q")
    assert np.max(np.abs(q[:n] - qd)) <= 1e-8
    assert abs(q[:n].sum()) <= 1e-9
    assert np.array_equal(q[n:N], q[owner[n:N]])          # ghost q follows its owner


def test_spmv_equals_dense(c0_oracle):
    s, o, xa, owner, N = c0_oracle
    n = s.n
    Hd = _dense_H(o, owner, n)
    rng = np.random.default_rng(1)
    xl = rng.normal(size=n)
    got = o.spmv(xl[owner])
    want = xl / o.get("Hdia_inv")[:n] + Hd @ xl
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))


def test_init_matvec_vectors(c0_oracle, ff):
    s, o, xa, owner, N = c0_oracle
    n = s.n
    el = np.array(s.elements)[s.type[:n] - 1]
    assert np.array_equal(o.get("b_s")[:n], np.array([-ff.atom(e)["chi"] for e in el]))
    assert np.all(o.get("b_t")[:n] == -1.0)
    assert np.array_equal(o.get("Hdia_inv")[:n], 1.0 / np.array([ff.eta(e) for e in el]))
