"""Writes the golden fixtures under tests/golden/ from the FP64 oracle.

The reference holds no expected outputs for this path, and its code cannot
be built or run here, so these vectors come from the oracle after the oracle
itself is pinned by tests/test_oracle_pins.py (closed forms, brute force,
dense solves).  They freeze the oracle's answers on inputs the water boxes do
not reach: carbon and nitrogen, the over-coordination (f1) branch of the
bond-order correction (O-O and C-C pairs), compressed bonds and isolated
molecules.  The GPU tests compare against the same files.

Cases (all non-periodic, QEq tol 1e-10 so the charges are solver-exact):
  tatb_cluster   the 384 atoms of data.tatb (TATB, C6H6N6O6 x 16) taken as
                 an isolated cluster in their file coordinates;
  compressed     O2 at 1.05 A, a compressed C-C-C chain, HCN, N2O4-like and
                 two waters 1.9 A apart;
  methanol_water a seeded 300-atom water/methanol cluster.

Run:  python tests/golden/gen_golden.py      (rewrites the .npz files)
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FFIELD = os.path.join(ROOT, "sw_reaxff_amd", "data", "ffield.reax")
TATB = os.path.join(ROOT, "sw_reaxff_amd", "data", "data.tatb")
TOL = 1e-10
PAD = 14.0


def read_lammps_charge_data(path):
    """Atoms section of a LAMMPS `atom_style charge` data file: id type q x y z."""
    with open(path) as f:
        lines = [ln.split("#")[0].strip() for ln in f]
    k = next(i for i, ln in enumerate(lines) if ln.startswith("Atoms"))
    rows = []
    for ln in lines[k + 1:]:
        if not ln:
            if rows:
                break
            continue
        t = ln.split()
        if not t[0].lstrip("-").isdigit():
            break
        rows.append((int(t[0]), int(t[1]), float(t[2]), float(t[3]), float(t[4]), float(t[5])))
    rows.sort()
    a = np.array(rows)
    return a[:, 1].astype(np.int32), a[:, 3:6].copy()


def case_tatb():
    types, x = read_lammps_charge_data(TATB)
    return dict(elements=np.array(["C", "H", "O", "N"]), type=types, x=x)


def case_compressed():
    mol = []
    # O2 compressed to 1.05 A
    mol += [("O", 0.0, 0.0, 0.0), ("O", 1.05, 0.0, 0.0)]
    # C-C-C chain at 1.22 A (over-coordinated carbons)
    mol += [("C", 6.0, 0.0, 0.0), ("C", 7.22, 0.0, 0.0), ("C", 8.44, 0.0, 0.0), ("H", 5.0, 0.3, 0.0)]
    # HCN
    mol += [("H", 0.0, 6.0, 0.0), ("C", 1.07, 6.0, 0.0), ("N", 2.22, 6.0, 0.0)]
    # N2O4-like: N-N 1.75, N-O 1.19
    mol += [("N", 6.0, 6.0, 0.0), ("N", 7.75, 6.0, 0.0), ("O", 5.4, 7.0, 0.0), ("O", 5.4, 5.0, 0.0),
            ("O", 8.35, 7.0, 0.0), ("O", 8.35, 5.0, 0.0)]
    # two waters with O-O 1.9 A and an H between them
    mol += [("O", 0.0, 0.0, 6.0), ("H", 0.95, 0.0, 6.0), ("H", -0.3, 0.9, 6.0),
            ("O", 1.9, 0.0, 6.2), ("H", 2.2, 0.9, 6.2), ("H", 2.2, -0.9, 6.2)]
    els = ["C", "H", "O", "N"]
    types = np.array([els.index(m[0]) + 1 for m in mol], dtype=np.int32)
    x = np.array([m[1:] for m in mol], dtype=np.float64)
    return dict(elements=np.array(els), type=types, x=x)


def case_methanol_water():
    from sw_reaxff_amd import systems
    s = systems.cluster(300, seed=12)
    return dict(elements=np.array(s.elements), type=s.type.astype(np.int32), x=s.x.copy())


def run_oracle(case, threads=4):
    from oracle import pyoracle
    x, types, els = case["x"], case["type"], [str(e) for e in case["elements"]]
    n = x.shape[0]
    o = pyoracle.Oracle(FFIELD, els, tol=TOL, threads=threads)
    o.set_atoms(n, x, types, np.arange(1, n + 1, dtype=np.int32), np.arange(n, dtype=np.int32))
    o.build_verlet(12.0)
    its = o.step()
    bi, be = o.get("bond_index"), o.get("bond_end")
    sel = np.concatenate([np.arange(bi[i], be[i]) for i in range(n)]) if n else np.zeros(0, int)
    bonds, bo = o.get("bonds"), o.get("bo_data")
    hi, he, Hindex = o.get("hbond_index"), o.get("hbond_end"), o.get("Hindex")
    hb = o.get("hbonds")
    hcount = np.zeros(n, np.int32)
    hnbr = []
    for i in range(n):
        if Hindex[i] >= 0 and Hindex[i] < hi.size and i < n:
            H = Hindex[i]
            seg = hb["nbr"][hi[H]:he[H]]
            hcount[i] = seg.size
            hnbr.append(seg)
    fi, fe = o.get("far_index"), o.get("far_end")
    out = dict(
        tol=np.float64(TOL), lo=x.min(0) - PAD, hi=x.max(0) + PAD,
        q=o.get("q")[:n], s=o.get("s")[:n], t=o.get("t")[:n],
        iters=np.array(its, dtype=np.int32),
        far_count=(fe - fi).astype(np.int32),
        bond_count=(be - bi).astype(np.int32),
        bond_nbr=bonds["nbr"][sel].astype(np.int32),
        BO=o.get("BO")[sel], BOpi=o.get("BOpi").reshape(-1, 2)[sel], BO_s=bo["BO_s"][sel],
        dBOp=bo["dBOp"][sel],
        C1dbo=bo["C1dbo"][sel], C2dbo=bo["C2dbo"][sel], C3dbo=bo["C3dbo"][sel],
        C1dbopi=bo["C1dbopi"][sel], C2dbopi=bo["C2dbopi"][sel], C3dbopi=bo["C3dbopi"][sel],
        C4dbopi=bo["C4dbopi"][sel],
        C1dbopi2=bo["C1dbopi2"][sel], C2dbopi2=bo["C2dbopi2"][sel], C3dbopi2=bo["C3dbopi2"][sel],
        C4dbopi2=bo["C4dbopi2"][sel],
        hbond_count=hcount,
        hbond_nbr=np.concatenate(hnbr).astype(np.int32) if hnbr else np.zeros(0, np.int32),
        H_numnbrs=o.get("H_numnbrs")[:n].astype(np.int32),
    )
    for k in ("Deltap", "Deltap_boc", "Delta", "Delta_e", "Delta_val", "vlpex", "nlp", "nlp_temp", "Clp",
              "Delta_lp", "dDelta_lp", "Delta_lp_temp", "dDelta_lp_temp"):
        out[k] = o.get(k)[:n]
    out["Delta_boc"] = o.get("bo_dboc").reshape(-1, 2)[:n, 1]
    return out


CASES = {"tatb_cluster": case_tatb, "compressed": case_compressed, "methanol_water": case_methanol_water}


def load(name):
    with np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def main():
    for name, fn in CASES.items():
        case = fn()
        res = run_oracle(case)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **case, **res)
        print(f"{name}: n={case['x'].shape[0]} bonds={res['bond_nbr'].size} hbonds={res['hbond_nbr'].size} "
              f"iters={tuple(res['iters'])} max|q|={np.abs(res['q']).max():.3f}")


if __name__ == "__main__":
    main()
