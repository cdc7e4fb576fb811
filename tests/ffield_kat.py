"""An independent reader of the ReaxFF force-field text format, for known-answer tests.

It reads the raw numbers of `ffield.reax` by position, following the published
format (van Duin's file layout, as LAMMPS' reaxc_ffield reads it).  It shares
no code with the oracle's or the library's readers, so a test that derives
its expected values from here also pins those readers:

  * atoms: 4 lines of 8 numbers after the element name;
  * bonds: 2 lines of 8 numbers after the two 1-based atom indices;
  * off-diagonal: 6 numbers after the two indices (values > 0 override the
    mixed radii r_s, r_p, r_pp).

Cited reference reader: reaxc_ffield_sunway.cpp:35-518 (eta is stored as
2x the file value at :189; the bond-order cutoff is 0.01 x gp[29] per
pair_reaxc_sunway.cpp:202-362 and reaxc_ffield_sunway.cpp).
"""
from __future__ import annotations

import math

ATOM_KEYS = [
    # line 1
    "r_s", "valency", "mass", "r_vdw", "epsilon", "gamma", "r_pi", "valency_e",
    # line 2
    "alpha", "gamma_w", "valency_boc", "p_ovun5", "_u0", "chi", "eta_file", "p_hbond",
    # line 3
    "r_pi_pi", "p_lp2", "_heat", "b_o_131", "b_o_132", "b_o_133", "_u1", "_u2",
    # line 4
    "p_ovun2", "p_val3", "_u3", "valency_val", "p_val5", "rcore2", "ecore2", "acore2",
]
BOND_KEYS = [
    "De_s", "De_p", "De_pp", "p_be1", "p_bo5", "v13cor", "p_bo6", "p_ovun1",
    "p_be2", "p_bo3", "p_bo4", "_u0", "p_bo1", "p_bo2", "ovc", "_u1",
]


class FField:
    def __init__(self, path: str):
        with open(path) as f:
            lines = f.read().splitlines()
        k = 1
        ngp = int(lines[k].split()[0])
        self.gp = [float(lines[k + 1 + i].split()[0]) for i in range(ngp)]
        k += 1 + ngp
        nat = int(lines[k].split()[0])
        k += 4
        self.atoms = {}
        self.order = []
        for _ in range(nat):
            t = lines[k].split()
            name, vals = t[0], [float(v) for v in t[1:9]]
            for m in range(1, 4):
                vals += [float(v) for v in lines[k + m].split()[:8]]
            self.atoms[name] = dict(zip(ATOM_KEYS, vals))
            self.order.append(name)
            k += 4
        nb = int(lines[k].split()[0])
        k += 2
        self.bonds = {}
        for _ in range(nb):
            t = lines[k].split()
            i, j = int(t[0]) - 1, int(t[1]) - 1
            vals = [float(v) for v in t[2:10]] + [float(v) for v in lines[k + 1].split()[:8]]
            d = dict(zip(BOND_KEYS, vals))
            self.bonds[(i, j)] = d
            self.bonds[(j, i)] = d
            k += 2
        no = int(lines[k].split()[0])
        k += 1
        self.offdiag = {}
        for _ in range(no):
            t = lines[k].split()
            i, j = int(t[0]) - 1, int(t[1]) - 1
            vals = [float(v) for v in t[2:8]]
            self.offdiag[(i, j)] = self.offdiag[(j, i)] = dict(zip(["D", "r_vdw", "alpha", "r_s", "r_p", "r_pp"], vals))
            k += 1
        # valence angles: end, centre, end (1-based) and theta_00, p_val1,
        # p_val2, p_coa1, p_val7, p_pen1, p_val4 (each line also serves the
        # reversed triple); the torsion block (a count line, then that many
        # lines); then the hydrogen bonds: acceptor, H, acceptor (1-based)
        # and r0_hb, p_hb1, p_hb2, p_hb3
        na = int(lines[k].split()[0])
        k += 1
        self.angles = {}
        for _ in range(na):
            t = lines[k].split()
            a, b, c = int(t[0]) - 1, int(t[1]) - 1, int(t[2]) - 1
            prm = dict(zip(["theta_00", "p_val1", "p_val2", "p_coa1", "p_val7", "p_pen1", "p_val4"],
                           [float(v) for v in t[3:10]]))
            self.angles.setdefault((a, b, c), []).append(prm)
            if c != a:
                self.angles.setdefault((c, b, a), []).append(prm)
            k += 1
        # torsions: "i j k l V1 V2 V3 p_tor1 p_cot1" (1-based); an explicit
        # line serves (i,j,k,l) and (l,k,j,i), a "0 j k 0" line every end
        # pair not given explicitly
        nt4 = int(lines[k].split()[0])
        k += 1
        explicit, wild = {}, {}
        for _ in range(nt4):
            t = lines[k].split()
            a, b, c, d = (int(v) - 1 for v in t[:4])
            prm = dict(zip(["V1", "V2", "V3", "p_tor1", "p_cot1"], [float(v) for v in t[4:9]]))
            if a >= 0 and d >= 0:
                explicit[(a, b, c, d)] = explicit[(d, c, b, a)] = prm
            else:
                wild[(b, c)] = wild[(c, b)] = prm
            k += 1
        self.torsions_explicit, self.torsions_wild = explicit, wild
        nh = int(lines[k].split()[0])
        k += 1
        self.hbonds = {}
        for _ in range(nh):
            t = lines[k].split()
            self.hbonds[(int(t[0]) - 1, int(t[1]) - 1, int(t[2]) - 1)] = dict(
                zip(["r0_hb", "p_hb1", "p_hb2", "p_hb3"], [float(v) for v in t[3:7]]))
            k += 1

    # --- derived quantities (standard ReaxFF definitions)
    @property
    def bo_cut(self) -> float:
        return 0.01 * self.gp[29]

    def idx(self, el: str) -> int:
        return self.order.index(el)

    def atom(self, el: str) -> dict:
        return self.atoms[el]

    def hbond(self, i: str, j: str, k: str):
        """Hydrogen-bond parameters of acceptor i, hydrogen j, acceptor k (None if absent)."""
        return self.hbonds.get((self.idx(i), self.idx(j), self.idx(k)))

    def angle(self, i: str, j: str, k: str):
        """Valence-angle parameter sets of end i, centre j, end k."""
        return self.angles.get((self.idx(i), self.idx(j), self.idx(k)), [])

    def torsion(self, i: str, j: str, k: str, l: str):
        """Torsion parameters of i-j-k-l (None if the quadruple has no term)."""
        key = (self.idx(i), self.idx(j), self.idx(k), self.idx(l))
        return self.torsions_explicit.get(key, self.torsions_wild.get(key[1:3]))

    def eta(self, el: str) -> float:
        return 2.0 * self.atoms[el]["eta_file"]

    def pair(self, a: str, b: str) -> dict:
        """Mixed bond radii and the bond-order parameters of the pair."""
        A, B = self.atoms[a], self.atoms[b]
        ia, ib = self.idx(a), self.idx(b)
        r_s = 0.5 * (A["r_s"] + B["r_s"])
        r_p = 0.5 * (A["r_pi"] + B["r_pi"])
        r_pp = 0.5 * (A["r_pi_pi"] + B["r_pi_pi"])
        od = self.offdiag.get((ia, ib))
        if od is not None:
            if od["r_s"] > 0:
                r_s = od["r_s"]
            if od["r_p"] > 0:
                r_p = od["r_p"]
            if od["r_pp"] > 0:
                r_pp = od["r_pp"]
        bd = self.bonds[(ia, ib)]
        return dict(bd, r_s=r_s, r_p=r_p, r_pp=r_pp,
                    p_boc3=math.sqrt(A["b_o_132"] * B["b_o_132"]),
                    p_boc4=math.sqrt(A["b_o_131"] * B["b_o_131"]),
                    p_boc5=math.sqrt(A["b_o_133"] * B["b_o_133"]))

    def bop(self, a: str, b: str, d: float):
        """Uncorrected bond-order terms (sigma, pi, pi-pi) at distance d,
        each gated on both atoms having the radius (ReaxFF BO' definition)."""
        A, B, t = self.atoms[a], self.atoms[b], self.pair(a, b)
        s = pi = pp = 0.0
        if A["r_s"] > 0 and B["r_s"] > 0:
            s = (1.0 + self.bo_cut) * math.exp(t["p_bo1"] * (d / t["r_s"]) ** t["p_bo2"])
        if A["r_pi"] > 0 and B["r_pi"] > 0:
            pi = math.exp(t["p_bo3"] * (d / t["r_p"]) ** t["p_bo4"])
        if A["r_pi_pi"] > 0 and B["r_pi_pi"] > 0:
            pp = math.exp(t["p_bo5"] * (d / t["r_pp"]) ** t["p_bo6"])
        return s, pi, pp

    def vdw_pair(self, a: str, b: str) -> dict:
        """Non-bonded pair parameters: geometric means of the atom values
        (r_vdW = 2 sqrt(r_vdw_a r_vdw_b)), the off-diagonal block's D, r_vdW
        (2x its value) and alpha when > 0, the Coulomb shielding
        (gamma_a gamma_b)^-1.5 and the inner-wall terms."""
        A, B = self.atoms[a], self.atoms[b]
        d = dict(D=math.sqrt(A["epsilon"] * B["epsilon"]), alpha=math.sqrt(A["alpha"] * B["alpha"]),
                 r_vdW=2.0 * math.sqrt(A["r_vdw"] * B["r_vdw"]), gamma_w=math.sqrt(A["gamma_w"] * B["gamma_w"]),
                 gamma=(A["gamma"] * B["gamma"]) ** -1.5, rcore=math.sqrt(A["rcore2"] * B["rcore2"]),
                 ecore=math.sqrt(A["ecore2"] * B["ecore2"]), acore=math.sqrt(A["acore2"] * B["acore2"]))
        od = self.offdiag.get((self.idx(a), self.idx(b)))
        if od is not None:
            if od["D"] > 0:
                d["D"] = od["D"]
            if od["r_vdw"] > 0:
                d["r_vdW"] = 2.0 * od["r_vdw"]
            if od["alpha"] > 0:
                d["alpha"] = od["alpha"]
        return d

    def vdw_type(self) -> int:
        """1 shielding, 2 inner wall, 3 both: the first element's setting."""
        A = self.atoms[self.order[0]]
        wall = A["rcore2"] > 0.01 and A["acore2"] > 0.01
        shield = A["gamma_w"] > 0.5
        return (3 if shield else 2) if wall else 1

    def nonbonded_pair_energy(self, a: str, b: str, r: float, qa: float, qb: float):
        """(E_vdW, E_Coulomb) of one pair at distance r: tapered shielded
        Morse (+ inner wall) and tapered shielded Coulomb (ReaxFF)."""
        P, p = self.vdw_pair(a, b), self.gp[28]
        vt = self.vdw_type()
        tap = self.taper_closed_form(r)
        if vt in (1, 3):
            f13 = (r ** p + P["gamma_w"] ** -p) ** (1.0 / p)
        else:
            f13 = r
        x = P["alpha"] * (1.0 - f13 / P["r_vdW"])
        ev = P["D"] * (math.exp(x) - 2.0 * math.exp(0.5 * x))
        if vt in (2, 3):
            ev += P["ecore"] * math.exp(P["acore"] * (1.0 - r / P["rcore"]))
        ec = 332.06371 * qa * qb / (r ** 3 + P["gamma"]) ** (1.0 / 3.0)
        return tap * ev, tap * ec

    def taper_closed_form(self, r: float) -> float:
        """7th-order taper for swa = 0: 20x^7 - 70x^6 + 84x^5 - 35x^4 + 1, x = r/swb."""
        x = r / self.gp[12]
        return 20 * x ** 7 - 70 * x ** 6 + 84 * x ** 5 - 35 * x ** 4 + 1

    def h_offdiag(self, a: str, b: str, r: float) -> float:
        """Shielded Coulomb H entry in kcal/mol-per-e^2 units of the QEq matrix."""
        g = (self.atoms[a]["gamma"] * self.atoms[b]["gamma"]) ** -1.5
        return self.taper_closed_form(r) * 14.4 / (r ** 3 + g) ** (1.0 / 3.0)
