"""ctypes binding of libswreax.so (the C-ABI declared in include/swreax.h).

The library is the product: every call here goes to HIP code on the GPU.
There is no CPU fallback — if the library is missing, or no GPU is visible
when a device call is made, these functions raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libswreax.so")

SWR_OK = 0
ERRORS = {1: "SWR_ERR_ARG", 2: "SWR_ERR_BOND_CAP", 3: "SWR_ERR_HBOND_CAP", 4: "SWR_ERR_NOMEM",
          5: "SWR_ERR_FFIELD", 6: "SWR_ERR_STATE", 7: "SWR_ERR_BIN", 8: "SWR_ERR_NUMERIC", 10: "SWR_ERR_HIP"}

# Every function the C-ABI header declares (checked by the CPU test suite).
EXPORTS = [
    "swr_control_default", "swr_last_error", "swr_create", "swr_destroy", "swr_set_atoms", "swr_set_capacities",
    "swr_npair_full_bin_ghost_build", "swr_write_reax_atoms_and_pack", "swr_reset", "swr_write_reax_lists",
    "swr_init_forces_noqeq_full", "swr_write_reax_lists_init_forces", "swr_validate_lists", "swr_bo",
    "swr_qeq_init_matvec", "swr_sparse_matvec", "swr_qeq_cg", "swr_calculate_q", "swr_qeq_pre_force",
    "swr_pair_compute", "swr_sync", "swr_get", "swr_put", "swr_get_scalar", "swr_device_ptr",
    "swr_lmp_open", "swr_lmp_close", "swr_lmp_command", "swr_lmp_create_atoms", "swr_lmp_set_history",
    "swr_lmp_setup", "swr_lmp_run",
    "swr_lmp_system", "swr_lmp_stat", "swr_ktimer_enable", "swr_ktimer_read", "swr_device_synchronize",
    "swr_sparse_matvec_h", "swr_npair_full_bin_atomonly_build", "swr_sort_bonds", "swr_set_batch",
    "swr_new_atoms", "swr_batch_pack", "swr_vdw_coulomb_energy", "swr_hist_save", "swr_hist_load",
    "swr_hist_free", "swr_atom_energy", "swr_add_dbond_forces", "swr_hydrogen_bonds",
    "swr_compute_forces", "swr_set_masses", "swr_valence_angles", "swr_torsion_angles", "swr_fix_nve_initial_integrate", "swr_fix_nve_final_integrate",
    "swr_autotune", "swr_autotune_load", "swr_set_shape", "swr_get_shape", "swr_tune_report",
    "swr_qeq_cg_noalloc",
]


class SwrError(RuntimeError):
    def __init__(self, code, where, msg=""):
        self.code = code
        super().__init__(f"{where}: {ERRORS.get(code, code)} {msg}".strip())


class Control(C.Structure):
    _fields_ = [("bond_cut", C.c_double), ("hbond_cut", C.c_double), ("thb_cut", C.c_double),
                ("safezone", C.c_double), ("saferzone", C.c_double), ("mincap", C.c_int), ("maxfar", C.c_int),
                ("swa", C.c_double), ("swb", C.c_double), ("tol", C.c_double), ("skin", C.c_double),
                ("qeq_mixed", C.c_int), ("qeq_imax", C.c_int), ("enobonds", C.c_int), ("thb_cutsq", C.c_double)]


_lib = None


def lib():
    """Loads libswreax.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                               "(make -C sw_reaxff_amd/csrc)")
        L = C.CDLL(LIB_PATH)
        vp, ip, dp, i64 = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_int64
        L.swr_control_default.argtypes = [C.POINTER(Control)]
        L.swr_last_error.restype = C.c_char_p
        L.swr_create.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_char_p), C.POINTER(Control), C.POINTER(vp)]
        L.swr_destroy.argtypes = [vp]
        L.swr_set_atoms.argtypes = [vp, C.c_int, C.c_int, ip, ip, dp, dp, ip, C.POINTER(C.c_int8)]
        L.swr_set_capacities.argtypes = [vp, ip, ip]
        L.swr_npair_full_bin_ghost_build.argtypes = [vp, C.c_double]
        L.swr_npair_full_bin_atomonly_build.argtypes = [vp, C.c_double]
        L.swr_set_batch.argtypes = [vp, C.c_int, ip]
        L.swr_new_atoms.argtypes = [vp]
        for nm in ["swr_write_reax_atoms_and_pack", "swr_reset", "swr_write_reax_lists", "swr_init_forces_noqeq_full",
                   "swr_write_reax_lists_init_forces", "swr_validate_lists", "swr_sort_bonds", "swr_bo",
                   "swr_qeq_init_matvec",
                   "swr_calculate_q", "swr_pair_compute", "swr_sync", "swr_vdw_coulomb_energy",
                   "swr_atom_energy", "swr_add_dbond_forces", "swr_hydrogen_bonds", "swr_compute_forces",
                   "swr_valence_angles", "swr_torsion_angles"]:
            getattr(L, nm).argtypes = [vp]
        L.swr_set_masses.argtypes = [vp, C.c_int, dp]
        L.swr_fix_nve_initial_integrate.argtypes = [vp, C.c_double, C.c_double]
        L.swr_fix_nve_final_integrate.argtypes = [vp, C.c_double, C.c_double]
        L.swr_sparse_matvec.argtypes = [vp, vp, vp]
        L.swr_sparse_matvec_h.argtypes = [vp, dp, dp]
        L.swr_qeq_cg.argtypes = [vp, ip, ip]
        L.swr_qeq_cg_noalloc.argtypes = [vp, vp, vp, ip, ip]
        L.swr_qeq_pre_force.argtypes = [vp, ip, ip]
        L.swr_get.restype = i64
        L.swr_get.argtypes = [vp, C.c_char_p, vp]
        L.swr_put.argtypes = [vp, C.c_char_p, vp, i64]
        L.swr_get_scalar.restype = i64
        L.swr_get_scalar.argtypes = [vp, C.c_char_p]
        L.swr_device_ptr.restype = vp
        L.swr_device_ptr.argtypes = [vp, C.c_char_p]
        L.swr_lmp_open.argtypes = [C.POINTER(vp)]
        L.swr_lmp_close.argtypes = [vp]
        L.swr_lmp_command.argtypes = [vp, C.c_char_p]
        L.swr_lmp_create_atoms.argtypes = [vp, C.c_int, dp, dp, C.c_int, dp, dp, ip, ip]
        L.swr_lmp_set_history.argtypes = [vp, C.c_int, C.c_void_p, C.c_void_p]
        L.swr_batch_pack.argtypes = [C.c_int] + [C.c_void_p] * 7 + [C.c_double, C.c_longlong] + [C.c_void_p] * 9
        L.swr_hist_save.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.POINTER(vp)]
        L.swr_hist_load.argtypes = [vp, C.c_int, C.c_int, vp]
        L.swr_hist_free.argtypes = [vp]
        L.swr_lmp_setup.argtypes = [vp]
        L.swr_lmp_run.argtypes = [vp, C.c_int]
        L.swr_lmp_system.restype = vp
        L.swr_lmp_system.argtypes = [vp]
        L.swr_lmp_stat.restype = C.c_double
        L.swr_lmp_stat.argtypes = [vp, C.c_char_p]
        L.swr_ktimer_enable.argtypes = [vp, C.c_int]
        L.swr_device_synchronize.argtypes = []
        L.swr_ktimer_read.argtypes = [vp, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_longlong), C.c_int]
        L.swr_autotune.argtypes = [vp, C.c_int]
        L.swr_autotune_load.argtypes = [vp]
        L.swr_set_shape.argtypes = [vp, C.c_char_p, C.c_char_p]
        L.swr_get_shape.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_int]
        L.swr_tune_report.restype = i64
        L.swr_tune_report.argtypes = [vp, C.c_char_p, i64]
        _lib = L
    return _lib


def check(rc, where):
    if rc != SWR_OK:
        raise SwrError(rc, where, lib().swr_last_error().decode(errors="replace"))


def ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


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

DTYPES = {
    "x": np.float64, "q": np.float64, "v": np.float64, "tag": np.int32, "type": np.int32, "owner": np.int32,
    "img": np.int8, "packed_atoms": PACK_DT, "verlet_offsets": np.int64, "verlet_list": np.int32, "far": FAR_DT,
    "far_index": np.int32, "far_end": np.int32, "num_bonds": np.int32, "num_hbonds": np.int32,
    "Hindex": np.int32, "bond_index": np.int32, "bond_end": np.int32, "bonds": BOND_DT, "bo_data": BO_DT,
    "hbond_index": np.int32, "hbond_end": np.int32, "hbonds": HBOND_DT, "hbond_far": np.int64,
    "H_val32": np.float32, "sys_iters": np.int32,
    "bond_mark": np.int32, "H_firstnbr": np.int32, "H_numnbrs": np.int32, "H_jlist": np.int32,
    "panel_rows": np.int32, "panel_begin": np.int32, "panel_count": np.int32, "panel_columns": np.int32,
    "H_jloc": np.uint16,
}
PANEL_MAXC = 3072   # columns a panel's list holds (SWR_PANEL_MAXC)


class System:
    """Low-level handle over one swr_system (one C-ABI call per reference stage)."""

    def __init__(self, ffield, elements, **ctl_kw):
        L = lib()
        self.L = L
        ctl = Control()
        L.swr_control_default(C.byref(ctl))
        for k, v in ctl_kw.items():
            setattr(ctl, k, v)
        arr = (C.c_char_p * len(elements))(*[e.encode() for e in elements])
        h = C.c_void_p()
        check(L.swr_create(ffield.encode(), len(elements), arr, C.byref(ctl), C.byref(h)), "swr_create")
        self.h = h
        self.owned = True

    @classmethod
    def wrap(cls, handle):
        obj = cls.__new__(cls)
        obj.L = lib()
        obj.h = C.c_void_p(handle)
        obj.owned = False
        return obj

    def close(self):
        if getattr(self, "h", None) and self.owned:
            self.L.swr_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_atoms(self, n, x, type_, tag, owner=None, img=None, q=None):
        N = int(x.shape[0])
        xx = np.ascontiguousarray(x, dtype=np.float64)
        tt = np.ascontiguousarray(type_, dtype=np.int32)
        gg = np.ascontiguousarray(tag, dtype=np.int32)
        oo = np.ascontiguousarray(owner if owner is not None else np.arange(N), dtype=np.int32)
        im = np.ascontiguousarray(img if img is not None else np.zeros((N, 3)), dtype=np.int8)
        qq = np.ascontiguousarray(q, dtype=np.float64) if q is not None else None   # None: zero charges
        check(self.L.swr_set_atoms(self.h, int(n), N, ptr(gg, C.c_int), ptr(tt, C.c_int), ptr(xx, C.c_double),
                                   ptr(qq, C.c_double) if qq is not None else None, ptr(oo, C.c_int),
                                   ptr(im, C.c_int8)), "swr_set_atoms")

    def set_capacities(self, num_bonds, num_hbonds):
        a = np.ascontiguousarray(num_bonds, dtype=np.int32)
        b = np.ascontiguousarray(num_hbonds, dtype=np.int32)
        check(self.L.swr_set_capacities(self.h, ptr(a, C.c_int), ptr(b, C.c_int)), "swr_set_capacities")

    def call(self, name, *args):
        check(getattr(self.L, "swr_" + name)(self.h, *args), "swr_" + name)

    def build_neighbors(self, cutneigh=0.0):
        self.call("npair_full_bin_ghost_build", C.c_double(cutneigh))

    def spmv(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        b = np.zeros(self.scalar("n"))
        check(self.L.swr_sparse_matvec_h(self.h, ptr(x, C.c_double), ptr(b, C.c_double)), "swr_sparse_matvec_h")
        return b

    def qeq_cg(self):
        a, b = C.c_int(0), C.c_int(0)
        check(self.L.swr_qeq_cg(self.h, C.byref(a), C.byref(b)), "swr_qeq_cg")
        return a.value, b.value

    def sync(self):
        check(self.L.swr_sync(self.h), "swr_sync")

    def get(self, name):
        nb = self.L.swr_get(self.h, name.encode(), None)
        if nb < 0:
            raise KeyError(name)
        dt = np.dtype(DTYPES.get(name, np.float64))
        out = np.empty(nb // dt.itemsize, dtype=dt)
        if nb:
            got = self.L.swr_get(self.h, name.encode(), out.ctypes.data_as(C.c_void_p))
            if got < 0:
                raise SwrError(10, "swr_get " + name, self.L.swr_last_error().decode())
        return out

    def put(self, name, arr):
        arr = np.ascontiguousarray(arr)
        check(self.L.swr_put(self.h, name.encode(), arr.ctypes.data_as(C.c_void_p), arr.nbytes), "swr_put " + name)

    def scalar(self, name):
        return self.L.swr_get_scalar(self.h, name.encode())

    # ---- autotuner of launch shapes
    def autotune(self, reps=5):
        """Times every launch shape of the tunable kernels on the current
        state and keeps the fastest; returns the report (dict)."""
        check(self.L.swr_autotune(self.h, int(reps)), "swr_autotune")
        return self.tune_report()

    def tune_report(self):
        import json
        nb = self.L.swr_tune_report(self.h, None, 0)
        buf = C.create_string_buffer(int(nb) + 1)
        self.L.swr_tune_report(self.h, buf, nb + 1)
        return json.loads(buf.value.decode()) if nb > 0 else {}

    def set_shape(self, family, shape):
        check(self.L.swr_set_shape(self.h, family.encode(), shape.encode()), f"swr_set_shape {family} {shape}")

    def shape(self, family):
        buf = C.create_string_buffer(32)
        check(self.L.swr_get_shape(self.h, family.encode(), buf, 32), f"swr_get_shape {family}")
        return buf.value.decode()

    def ktimer_enable(self, on=True):
        check(self.L.swr_ktimer_enable(self.h, 1 if on else 0), "swr_ktimer_enable")

    def ktimer_read(self, slot, reset=False):
        ms, cnt = C.c_double(0), C.c_longlong(0)
        check(self.L.swr_ktimer_read(self.h, slot.encode(), C.byref(ms), C.byref(cnt), 1 if reset else 0),
              "swr_ktimer_read " + slot)
        return ms.value, cnt.value
