"""The mini-LAMMPS harness (C++ in libswreax.so) driven from Python.

    lmp = Lammps()
    lmp.commands('''
        pair_style reax/c/sunway NULL
        pair_coeff * * ffield.reax H O
        fix 2 all qeq/reax/sunway 1 0.0 10.0 1.0e-6 reax/c
        neighbor 2.0 bin
        neigh_modify every 10
        timestep 0.25
    ''')
    lmp.create_atoms(system)      # a sw_reaxff_amd.systems.System
    lmp.run(100)

Command names and arguments are those of `pair_style reax/c/sunway` and
`fix qeq/reax/sunway` (PairStyle pair_reaxc_sunway.h:26, settings
pair_reaxc_sunway.cpp:202-362; FixStyle fix_qeq_reax_sunway.h:26, arguments
fix_qeq_reax_sunway.cpp:76-99); `reax/c` and `qeq/reax` resolve to them as
under `-sf sunway`.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import System, check, lib, ptr

DEFAULT_FFIELD = os.path.join(_lib.PKG_DIR, "data", "ffield.reax")


class Lammps:
    def __init__(self):
        self.L = lib()
        h = C.c_void_p()
        check(self.L.swr_lmp_open(C.byref(h)), "swr_lmp_open")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.swr_lmp_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def command(self, cmd: str):
        check(self.L.swr_lmp_command(self.h, cmd.encode()), cmd)

    def commands(self, text: str):
        for line in text.strip().splitlines():
            line = line.strip()
            if line:
                self.command(line)

    def create_atoms(self, s):
        x = np.ascontiguousarray(s.x, dtype=np.float64)
        v = np.ascontiguousarray(s.v, dtype=np.float64)
        t = np.ascontiguousarray(s.type, dtype=np.int32)
        g = np.ascontiguousarray(s.tag, dtype=np.int32)
        lo = np.ascontiguousarray(s.lo, dtype=np.float64)
        hi = np.ascontiguousarray(s.hi, dtype=np.float64)
        check(self.L.swr_lmp_create_atoms(self.h, s.n, ptr(lo, C.c_double), ptr(hi, C.c_double),
                                          1 if s.periodic else 0, ptr(x, C.c_double), ptr(v, C.c_double),
                                          ptr(t, C.c_int), ptr(g, C.c_int)), "create_atoms")

    def set_history(self, s_hist, t_hist):
        """QEq s/t histories (n x 5 each) that warm-start the setup solve."""
        sh = np.ascontiguousarray(s_hist, dtype=np.float64)
        th = np.ascontiguousarray(t_hist, dtype=np.float64)
        assert sh.shape == th.shape and sh.ndim == 2 and sh.shape[1] == 5
        check(self.L.swr_lmp_set_history(self.h, sh.shape[0], sh.ctypes.data, th.ctypes.data), "set_history")

    def setup(self):
        check(self.L.swr_lmp_setup(self.h), "setup")

    def run(self, nsteps: int):
        check(self.L.swr_lmp_run(self.h, int(nsteps)), f"run {nsteps}")

    def stat(self, name: str) -> float:
        return self.L.swr_lmp_stat(self.h, name.encode())

    @property
    def system(self) -> System:
        return System.wrap(self.L.swr_lmp_system(self.h))


def standard_script(elements, ffield=DEFAULT_FFIELD, skin=2.0, every=10, dt=0.25, tol=1e-6, mixed=False,
                    profile=False, forces=False):
    """The LAMMPS input this path is benchmarked with (pair_style NULL defaults).

    forces=True adds the harness command `forces yes`: the energy terms built
    here and NVE with their forces (without it NVE moves atoms at constant
    velocity and only the path runs)."""
    return "\n".join([
        "units real",
        "atom_style charge",
        "pair_style reax/c/sunway NULL",
        f"pair_coeff * * {ffield} {' '.join(elements)}",
        f"fix 2 all qeq/reax/sunway 1 0.0 10.0 {tol:g} reax/c" + (" mixed" if mixed else ""),
        f"neighbor {skin} bin",
        f"neigh_modify every {every}",
        f"timestep {dt}",
        f"profile {'yes' if profile else 'no'}",
    ] + (["forces yes"] if forces else []))
