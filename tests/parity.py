"""Shared helpers for the GPU-vs-oracle parity tests.

The GPU side runs through the C-ABI (libswreax.so); the oracle side is the
CPU FP64 restatement (oracle/).  Both see identical atoms: the GPU harness
builds the ghost images and the Verlet list, those arrays are downloaded,
and the oracle is fed the same atoms and (for stage comparisons) the same
Verlet order, so that summation orders agree.
"""
from __future__ import annotations

from oracle.compare import (BO_FIELDS, BO_RTOL, WS_ARRAYS, atoms_of, compare_energy_forces,  # noqa: F401
                            compare_state, feed_oracle, gpu_energy_forces, history_of, oracle_energy_forces,
                            ranges, rel_err, verlet_of)
from sw_reaxff_amd import systems
from sw_reaxff_amd.lammps import Lammps, standard_script

# Parity tolerance of the charges at the reference tolerance: 1e-6 e (lists
# bit-exact as sets and bond orders 1e-10 relative: BO_RTOL).
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


def per_atom(index, end, arr, i):
    return arr[index[i]:end[i]]


def make_oracle(oracle_lib, ffield, elements, A, verlet=None, threads=8, tol=1e-6):
    o = oracle_lib.Oracle(ffield, elements, threads=threads, tol=tol)
    o.set_atoms(A["n"], A["x"], A["type"], A["tag"], A["owner"], q=A["q"])
    if verlet is not None:
        o.set_verlet(*verlet)
    return o


def water_c0(seed=1):
    return systems.water_box(648, seed=seed)
