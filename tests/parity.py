"""Shared helpers for the GPU-vs-oracle parity tests.

The GPU side runs through the C-ABI (libswreax.so); the oracle side is the
CPU FP64 restatement (oracle/).  Both see identical atoms: the GPU harness
builds the ghost images and the Verlet list, those arrays are downloaded,
and the oracle is fed the same atoms and (for stage comparisons) the same
Verlet order, so that summation orders agree.
"""
from __future__ import annotations

import numpy as np

from sw_reaxff_amd import systems
from sw_reaxff_amd.lammps import Lammps, standard_script

# Parity tolerances of the path: lists bit-exact as sets, bond orders 1e-10
# relative, charges 1e-6 e.
BO_RTOL = 1e-10
Q_ATOL = 1e-6


def harness(system, ffield, steps=0, every=10, dt=0.25, mixed=False, profile=False, tol=1e-6, forces=False):
    lmp = Lammps()
    lmp.commands(standard_script(system.elements, ffield=ffield, every=every, dt=dt, mixed=mixed,
                                 profile=profile, tol=tol, forces=forces))
    lmp.create_atoms(system)
    lmp.setup()
    if steps:
        lmp.run(steps)
    return lmp


def atoms_of(sysh):
    x = sysh.get("x").reshape(-1, 3)
    return dict(n=sysh.scalar("n"), N=sysh.scalar("N"), x=x, type=sysh.get("type"), tag=sysh.get("tag"),
                owner=sysh.get("owner"), q=sysh.get("q"))


def verlet_of(sysh):
    off = sysh.get("verlet_offsets")
    lst = sysh.get("verlet_list")
    return off, lst[: off[-1]]


def per_atom(index, end, arr, i):
    return arr[index[i]:end[i]]


def make_oracle(oracle_lib, ffield, elements, A, verlet=None, threads=8, tol=1e-6):
    o = oracle_lib.Oracle(ffield, elements, threads=threads, tol=tol)
    o.set_atoms(A["n"], A["x"], A["type"], A["tag"], A["owner"], q=A["q"])
    if verlet is not None:
        o.set_verlet(*verlet)
    return o


def rel_err(a, b, floor=1e-300):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.maximum(np.abs(b), floor)
    return np.max(np.abs(a - b) / den) if a.size else 0.0


def water_c0(seed=1):
    return systems.water_box(648, seed=seed)
