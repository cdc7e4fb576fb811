"""ctypes binding of the CPU FP64 oracle (test infrastructure only).

Only the test suite, the smoke check and the benchmark's CPU-baseline leg
import this module.  It builds nothing: `make -C oracle` (or the package's
build step) produces liboracle_swreax.so next to this file.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_swreax.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_char_p), C.c_double, C.c_double,
                                 C.c_double, C.c_double, C.c_double, C.c_char_p, C.c_int]
        L.orc_destroy.argtypes = [vp]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_gp.restype = C.c_double
        L.orc_gp.argtypes = [vp, C.c_int]
        L.orc_sbp.argtypes = [vp, C.c_int, dp]
        L.orc_tbp.argtypes = [vp, C.c_int, C.c_int, dp]
        L.orc_control.restype = C.c_double
        L.orc_control.argtypes = [vp, C.c_char_p]
        L.orc_taper.argtypes = [vp, dp]
        L.orc_set_atoms.argtypes = [vp, C.c_int, C.c_int, ip, ip, ip, dp, dp]
        L.orc_set_q.argtypes = [vp, dp]
        L.orc_build_verlet.restype = C.c_int64
        L.orc_build_verlet.argtypes = [vp, C.c_double, C.c_int]
        L.orc_set_verlet.argtypes = [vp, C.POINTER(C.c_int64), ip]
        L.orc_write_reax_lists.restype = C.c_int64
        for nm in ["orc_write_reax_atoms", "orc_estimate_storages", "orc_reset", "orc_init_forces",
                   "orc_validate_lists", "orc_bo", "orc_qeq_init_matvec", "orc_qeq_calculate_q"]:
            getattr(L, nm).argtypes = [vp]
        L.orc_write_reax_lists.argtypes = [vp]
        L.orc_qeq_cg.argtypes = [vp, C.c_int]
        L.orc_set_cg_imax.argtypes = [vp, C.c_int]
        L.orc_qeq_pre_force.argtypes = [vp, ip, ip]
        L.orc_step.argtypes = [vp, ip, ip]
        L.orc_sparse_matvec.argtypes = [vp, dp, dp]
        L.orc_vdw_coulomb.argtypes = [vp, dp]
        L.orc_atom_energy.argtypes = [vp, C.c_int, dp]
        L.orc_dbond_forces.argtypes = [vp]
        L.orc_hydrogen_bonds.restype = C.c_double
        L.orc_hydrogen_bonds.argtypes = [vp]
        L.orc_valence_angles.argtypes = [vp, dp]
        L.orc_torsion_angles.argtypes = [vp, dp]
        L.orc_get_array.restype = C.c_int64
        L.orc_get_array.argtypes = [vp, C.c_char_p, vp]
        L.orc_set_array.argtypes = [vp, C.c_char_p, vp, C.c_int64]
        L.orc_get_scalar_i.restype = C.c_int64
        L.orc_get_scalar_i.argtypes = [vp, C.c_char_p]
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


# numpy dtypes of the layout structs (match include/swreax_layout.h)
FAR_DT = np.dtype([("nbr", "<i4"), ("type", "<i4"), ("orig_id", "<i4"), ("rel_box", "<i4", 3),
                   ("d", "<f8"), ("q", "<f8"), ("dvec", "<f8", 3)])
BOND_DT = np.dtype([("d", "<f8"), ("dvec", "<f8", 3), ("nbr", "<i4"), ("dbond_index", "<i4"),
                    ("rel_box", "<i4", 3), ("sym_index", "<i4")])
BO_DT = np.dtype([("BO_s", "<f8"), ("dBOp", "<f8", 3), ("dln_BOp_s", "<f8", 3), ("dln_BOp_pi", "<f8", 3),
                  ("dln_BOp_pi2", "<f8", 3), ("C1dbo", "<f8"), ("C2dbo", "<f8"), ("C3dbo", "<f8"),
                  ("C1dbopi", "<f8"), ("C2dbopi", "<f8"), ("C3dbopi", "<f8"), ("C4dbopi", "<f8"),
                  ("C1dbopi2", "<f8"), ("C2dbopi2", "<f8"), ("C3dbopi2", "<f8"), ("C4dbopi2", "<f8"),
                  ("CdboReduction", "<u8")])
HBOND_DT = np.dtype([("nbr", "<i4"), ("scl", "<i4"), ("ptr", "<u8")])
PACK_DT = np.dtype([("orig_id", "<i4"), ("type", "<i4"), ("x", "<f8", 3), ("q", "<f8")])
assert FAR_DT.itemsize == 64 and BOND_DT.itemsize == 56 and BO_DT.itemsize == 200
assert HBOND_DT.itemsize == 16 and PACK_DT.itemsize == 40

_DTYPES = {
    "verlet_offsets": np.int64, "verlet_list": np.int32, "packed_atoms": PACK_DT,
    "far_index": np.int64, "far_end": np.int64, "far": FAR_DT,
    "num_bonds": np.int32, "num_hbonds": np.int32, "Hindex": np.int32,
    "bond_index": np.int32, "bond_end": np.int32, "bonds": BOND_DT, "bo_data": BO_DT,
    "hbond_index": np.int32, "hbond_end": np.int32, "hbonds": HBOND_DT, "hbond_far": np.int64,
    "bond_mark": np.int32, "H_firstnbr": np.int32, "H_numnbrs": np.int32, "H_jlist": np.int32,
}


class Oracle:
    """One oracle system (mirrors the reference's pair + fix state)."""

    def __init__(self, ffield: str, elements, bond_cut=5.0, hbond_cut=7.5, swa=0.0, swb=10.0,
                 tol=1e-6, threads: int | None = None):
        L = lib()
        if threads is not None:
            L.orc_set_threads(int(threads))
        arr = (C.c_char_p * len(elements))(*[e.encode() for e in elements])
        err = C.create_string_buffer(512)
        self.h = L.orc_create(ffield.encode(), len(elements), arr, bond_cut, hbond_cut, swa, swb, tol,
                              err, 512)
        if not self.h:
            raise RuntimeError("oracle create failed: " + err.value.decode())
        self.L = L

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    # --- queries
    def gp(self, k):
        return self.L.orc_gp(self.h, k)

    def sbp(self, e):
        o = np.zeros(16)
        self.L.orc_sbp(self.h, e, _p(o, C.c_double))
        keys = ["r_s", "valency", "mass", "gamma", "r_pi", "valency_e", "nlp_opt", "valency_boc", "chi",
                "eta", "p_hbond", "r_pi_pi", "b_o_131", "b_o_132", "b_o_133", "valency_val"]
        return dict(zip(keys, o))

    def tbp(self, i, j):
        o = np.zeros(14)
        self.L.orc_tbp(self.h, i, j, _p(o, C.c_double))
        keys = ["p_bo1", "p_bo2", "p_bo3", "p_bo4", "p_bo5", "p_bo6", "r_s", "r_p", "r_pp", "p_boc3",
                "p_boc4", "p_boc5", "v13cor", "ovc"]
        return dict(zip(keys, o))

    def control(self, key):
        return self.L.orc_control(self.h, key.encode())

    def taper(self):
        o = np.zeros(8)
        self.L.orc_taper(self.h, _p(o, C.c_double))
        return o

    # --- atoms
    def set_atoms(self, n, x, type_, tag, owner, q=None):
        N = x.shape[0]
        self._keep = [np.ascontiguousarray(a) for a in (x.astype(np.float64), type_.astype(np.int32),
                                                        tag.astype(np.int32), owner.astype(np.int32))]
        xx, tt, gg, oo = self._keep
        qq = np.zeros(N) if q is None else np.ascontiguousarray(q, dtype=np.float64)
        rc = self.L.orc_set_atoms(self.h, int(n), int(N), _p(gg, C.c_int), _p(tt, C.c_int), _p(oo, C.c_int),
                                  _p(xx, C.c_double), _p(qq, C.c_double))
        if rc:
            raise RuntimeError("orc_set_atoms failed")

    def set_q(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        self.L.orc_set_q(self.h, _p(q, C.c_double))

    # --- stages
    def build_verlet(self, cut, brute=False):
        return self.L.orc_build_verlet(self.h, cut, 1 if brute else 0)

    def set_verlet(self, offsets, lst):
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        ls = np.ascontiguousarray(lst, dtype=np.int32)
        self.L.orc_set_verlet(self.h, _p(off, C.c_int64), _p(ls, C.c_int))

    def stage(self, name, *args):
        rc = getattr(self.L, "orc_" + name)(self.h, *args)
        return rc

    def step(self):
        a, b = C.c_int(0), C.c_int(0)
        rc = self.L.orc_step(self.h, C.byref(a), C.byref(b))
        if rc:
            raise RuntimeError(f"oracle step failed rc={rc}")
        return a.value, b.value

    def set_cg_imax(self, imax):
        """CG_v2's iteration cap (200 in the reference)."""
        self.L.orc_set_cg_imax(self.h, int(imax))

    def cg(self, which=0):
        """One CG_v2 solve (0: s with b_s, 1: t with b_t) from the current
        s or t; returns the reference's loop counter i."""
        return self.L.orc_qeq_cg(self.h, int(which))

    def pre_force(self):
        a, b = C.c_int(0), C.c_int(0)
        self.L.orc_qeq_pre_force(self.h, C.byref(a), C.byref(b))
        return a.value, b.value

    def vdw_coulomb(self):
        """Non-bonded pass over the current far list and packed charges: adds
        the gradient into fCdDelta; returns (eng_vdwl, eng_coul, virial[6])."""
        o = np.zeros(8)
        self.L.orc_vdw_coulomb(self.h, _p(o, C.c_double))
        return o[0], o[1], o[2:].copy()

    def atom_energy(self, enobonds=1):
        """Atom_Energy (bond, lone-pair, over/under-coordination, polarization
        energies) over the current bonds: accumulates Cdbo/Cdbopi/Cdbopi2 and
        CdDelta (fCdDelta[:, 3]); returns (e_bond, e_lp, e_ov, e_un, e_pol)."""
        o = np.zeros(5)
        self.L.orc_atom_energy(self.h, int(enobonds), _p(o, C.c_double))
        return o

    def hydrogen_bonds(self):
        """Hydrogen_Bonds over the current bonds and hbond lists: adds CEhb1 to
        Cdbo and the angle/distance terms to fCdDelta[:, 0..2]; returns e_hb."""
        return float(self.L.orc_hydrogen_bonds(self.h))

    def valence_angles(self):
        """Valence part of Merge_Torsion_Valence_Angles over the current bonds:
        adds to Cdbo/Cdbopi/Cdbopi2, CdDelta and fCdDelta[:, 0..2]; returns
        (e_ang, e_pen, e_coa)."""
        o = np.zeros(3)
        self.L.orc_valence_angles(self.h, _p(o, C.c_double))
        return o

    def torsion_angles(self):
        """Torsion part of Merge_Torsion_Valence_Angles over the current
        bonds: adds to Cdbo/Cdbopi, CdDelta and fCdDelta[:, 0..2]; returns
        (e_tor, e_con)."""
        o = np.zeros(2)
        self.L.orc_torsion_angles(self.h, _p(o, C.c_double))
        return o

    def dbond_forces(self):
        """Add_All_dBond_to_Forces: adds the bond-order chain rule to
        fCdDelta[:, 0..2] over all N atoms."""
        self.L.orc_dbond_forces(self.h)

    def spmv(self, xN):
        xN = np.ascontiguousarray(xN, dtype=np.float64)
        b = np.zeros(self.scalar("n"))
        self.L.orc_sparse_matvec(self.h, _p(xN, C.c_double), _p(b, C.c_double))
        return b

    def get(self, name):
        nb = self.L.orc_get_array(self.h, name.encode(), None)
        if nb < 0:
            raise KeyError(name)
        dt = np.dtype(_DTYPES.get(name, np.float64))
        out = np.empty(nb // dt.itemsize, dtype=dt)
        if nb:
            self.L.orc_get_array(self.h, name.encode(), out.ctypes.data_as(C.c_void_p))
        return out

    def set(self, name, arr):
        arr = np.ascontiguousarray(arr)
        rc = self.L.orc_set_array(self.h, name.encode(), arr.ctypes.data_as(C.c_void_p), arr.nbytes)
        if rc:
            raise KeyError(name)

    def scalar(self, name):
        return self.L.orc_get_scalar_i(self.h, name.encode())


# ---------------------------------------------------------------- forces, NVE
# Plain numpy restatements; pinned in tests/test_md.py by finite differences
# of the periodic energy (forces), a constant-force closed form and time
# reversibility (NVE).

def forces(fcd, owner, n):
    """Forces on the n local atoms: PairReaxCSunway::compute adds -fCdDelta
    to atom->f of every atom, ghosts included (pair_reaxc_sunway.cpp:746-760),
    and LAMMPS' reverse communication adds each ghost's force onto the local
    atom it images (owner[g])."""
    g = np.asarray(fcd, np.float64).reshape(-1, 4)[:, :3]
    N = owner.shape[0]
    f = -g[:n].copy()
    np.add.at(f, np.asarray(owner[n:N], np.int64), -g[n:N])
    return f


def nve_initial(x, v, f, mass, dtv, dtf):
    """fix_nve_initial_integrate_para, per-type mass branch
    (fix_nve_sw64.c:62-99): v += dtf/m f; x += dtv v.  mass is per atom."""
    dtfm = (dtf / np.asarray(mass, np.float64))[:, None]
    v = v + dtfm * f
    return x + dtv * v, v


def nve_final(v, f, mass, dtf):
    """fix_nve_final_integrate_para, per-type mass branch
    (fix_nve_sw64.c:138-171): v += dtf/m f."""
    return v + (dtf / np.asarray(mass, np.float64))[:, None] * f


FTM2V = 1.0 / 48.88821291 / 48.88821291   # force->ftm2v, LAMMPS real units
