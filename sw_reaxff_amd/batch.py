"""Batched evaluation of independent systems in one swr_system.

Many small systems (clusters, small periodic boxes) are too small to fill an
MI355X one at a time: a 500-atom cluster is a few hundred wavefronts per
stage, and the per-stage launch and host round trips dominate.  A Batch
packs K systems into one device system:

  * each system (with its ghost images when periodic) is translated into its
    own region of one virtual box, with regions at least the neighbour cutoff
    apart, so no pair crosses between systems and every list stage (Verlet,
    far, bonds, hbonds, BO) runs unchanged, once, over all atoms;
  * local atoms of all systems come first (system by system), then their
    ghosts; owner[] points at the system's own local atoms;
  * swr_set_batch gives the library the per-system row offsets, so QEq keeps
    per-system CG scalars, stopping tests and charge neutrality (each
    system's charges are those it gets alone).

Element types of the systems are mapped onto one combined element list.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib, systems
from .lammps import DEFAULT_FFIELD

CUTGHOST = 12.0          # max(nonb 10, hbond 7.5, 2 x bond 5) + 2 A skin
ELEMENTS = ["H", "O", "C", "N"]


@dataclass
class Packed:
    x: np.ndarray
    type: np.ndarray
    tag: np.ndarray
    owner: np.ndarray
    img: np.ndarray
    n: int
    row_off: np.ndarray      # [K+1] local-row offsets
    ghost_off: np.ndarray    # [K+1] ghost offsets (relative to n)
    shift: np.ndarray        # [K,3] translation applied to each system


def ghosts_of(s: systems.System):
    """Local + ghost coordinates, owner and periodic image of one system."""
    if not s.periodic:
        n = s.n
        return s.x.copy(), np.arange(n, dtype=np.int32), np.zeros((n, 3), np.int8)
    xa, owner = systems.borders(s.x, s.lo, s.hi, CUTGHOST)
    L = s.hi - s.lo
    img = np.rint((xa - s.x[owner]) / L).astype(np.int8)
    return xa, owner.astype(np.int32), img


def pack(batch: list[systems.System], elements=ELEMENTS) -> Packed:
    """Places the systems in disjoint regions of one box (see module doc)."""
    K = len(batch)
    parts = [ghosts_of(s) for s in batch]
    ext = []
    for s, (xa, _, _) in zip(batch, parts):
        lo = np.minimum(xa.min(0), s.lo) if s.n else s.lo
        hi = np.maximum(xa.max(0), s.hi) if s.n else s.hi
        ext.append((lo, hi))
    # uniform cubic cells on an m x m x m grid, each a system's extent plus the cutoff gap
    size = max(float(np.max(hi - lo)) for lo, hi in ext) + CUTGHOST + 1.0
    m = max(1, int(np.ceil(K ** (1.0 / 3.0) - 1e-9)))
    shifts = np.zeros((K, 3))
    for k, (lo, hi) in enumerate(ext):
        cell = np.array([k % m, (k // m) % m, k // (m * m)], dtype=np.float64)
        shifts[k] = cell * size - lo
    n = sum(s.n for s in batch)
    row_off = np.zeros(K + 1, np.int32)
    row_off[1:] = np.cumsum([s.n for s in batch])
    ng = [xa.shape[0] - s.n for s, (xa, _, _) in zip(batch, parts)]
    ghost_off = np.zeros(K + 1, np.int32)
    ghost_off[1:] = np.cumsum(ng)
    N = n + int(ghost_off[-1])
    x = np.empty((N, 3))
    typ = np.empty(N, np.int32)
    tag = np.empty(N, np.int32)
    owner = np.empty(N, np.int32)
    img = np.zeros((N, 3), np.int8)
    for k, (s, (xa, own, im)) in enumerate(zip(batch, parts)):
        tmap = np.array([elements.index(e) + 1 for e in s.elements], dtype=np.int32)
        a, b = row_off[k], row_off[k + 1]
        g0, g1 = n + ghost_off[k], n + ghost_off[k + 1]
        t_all = tmap[s.type[own] - 1]
        x[a:b] = xa[:s.n] + shifts[k]
        x[g0:g1] = xa[s.n:] + shifts[k]
        typ[a:b], typ[g0:g1] = t_all[:s.n], t_all[s.n:]
        tag[a:b] = s.tag + a
        tag[g0:g1] = s.tag[own[s.n:]] + a
        owner[a:b] = np.arange(a, b)
        owner[g0:g1] = own[s.n:] + a
        img[g0:g1] = im[s.n:]
    return Packed(x=x, type=typ, tag=tag, owner=owner, img=img, n=n, row_off=row_off, ghost_off=ghost_off,
                  shift=shifts)


class Batch:
    """One device system holding a batch; evaluate() runs nbr + QEq + lists + BO once for all."""

    def __init__(self, ffield=DEFAULT_FFIELD, elements=ELEMENTS, tol=1e-6, mixed=False):
        self.elements = list(elements)
        self.sys = _lib.System(ffield, self.elements, tol=tol, qeq_mixed=1 if mixed else 0)
        self.packed = None

    def load(self, batch: list[systems.System], q=None):
        """New systems: QEq starts from zero history and list capacities are estimated afresh."""
        P = pack(batch, self.elements)
        h = self.sys
        h.set_atoms(P.n, P.x, P.type, P.tag, owner=P.owner, img=P.img, q=q)
        _lib.check(h.L.swr_new_atoms(h.h), "swr_new_atoms")
        off = np.ascontiguousarray(P.row_off, dtype=np.int32)
        _lib.check(h.L.swr_set_batch(h.h, len(batch), off.ctypes.data_as(C.POINTER(C.c_int))), "swr_set_batch")
        self.packed = P
        return P

    def evaluate(self):
        """Neighbour list, QEq (per-system), far/bond/hbond lists, BO; returns max CG iterations (s, t)."""
        h = self.sys
        _lib.check(h.L.swr_npair_full_bin_ghost_build(h.h, C.c_double(CUTGHOST)), "swr_npair_full_bin_ghost_build")
        a, b = C.c_int(0), C.c_int(0)
        _lib.check(h.L.swr_qeq_pre_force(h.h, C.byref(a), C.byref(b)), "swr_qeq_pre_force")
        _lib.check(h.L.swr_pair_compute(h.h), "swr_pair_compute")
        return a.value, b.value

    def sync(self):
        self.sys.sync()

    # ---- per-system views of the results
    def rows(self, k):
        P = self.packed
        return slice(int(P.row_off[k]), int(P.row_off[k + 1]))

    def charges(self):
        q = self.sys.get("q")
        P = self.packed
        return [q[P.row_off[k]:P.row_off[k + 1]] for k in range(len(P.row_off) - 1)]

    def iterations(self):
        it = self.sys.get("sys_iters")
        return it.reshape(-1, 2)
