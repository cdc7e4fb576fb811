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

# Ghost shell and Verlet cutoff of a one-shot evaluation: max(nonb 10,
# hbond 7.5, 2 x bond 5).  The lists are built fresh for every evaluation,
# so no skin is needed (an MD run adds one so lists survive several steps);
# every result of a local atom (its far/bond/hbond lists, H row, corrected
# bond orders, which reach atoms within 2 x bond cutoff) lies inside it.
CUTGHOST = 10.0
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


def _buffer(bufs, name, shape, dtype):
    """A reusable output array of at least `shape` (first axis grows)."""
    if bufs is None:
        return np.empty(shape, dtype)
    b = bufs.get(name)
    if b is None or b.shape[0] < shape[0]:
        b = np.empty((max(shape[0], int(shape[0] * 1.25) + 16),) + tuple(shape[1:]), dtype)
        bufs[name] = b
    return b[:shape[0]]


def pack(batch: list[systems.System], elements=ELEMENTS, cutghost: float = CUTGHOST, bufs=None) -> Packed:
    """Places the systems in disjoint regions of one box (see module doc);
    ghosts, translation and numbering by swr_batch_pack (host C++).  `bufs`
    (a dict kept by the caller) lets successive packs reuse their output
    arrays; the returned arrays are then views into them."""
    L = _lib.lib()
    K = len(batch)
    nat = np.array([s.n for s in batch], dtype=np.int32)
    n = int(nat.sum())
    x = np.ascontiguousarray(np.concatenate([s.x for s in batch]) if n else np.zeros((0, 3)), dtype=np.float64)
    typ = np.empty(n, np.int32)
    tag = np.empty(n, np.int32)
    a = 0
    for s in batch:
        tmap = np.array([elements.index(e) + 1 for e in s.elements], dtype=np.int32)
        typ[a:a + s.n] = tmap[s.type - 1]
        tag[a:a + s.n] = s.tag
        a += s.n
    lo = np.ascontiguousarray([s.lo for s in batch], dtype=np.float64)
    hi = np.ascontiguousarray([s.hi for s in batch], dtype=np.float64)
    per = np.array([1 if s.periodic else 0 for s in batch], dtype=np.int32)
    N = C.c_longlong(0)
    args = (K, nat.ctypes.data, x.ctypes.data, typ.ctypes.data, tag.ctypes.data, lo.ctypes.data, hi.ctypes.data,
            per.ctypes.data, C.c_double(cutghost))
    _lib.check(L.swr_batch_pack(*args, 0, None, None, None, None, None, None, None, None, C.byref(N)),
               "swr_batch_pack")
    Nv = N.value
    xo = _buffer(bufs, "x", (Nv, 3), np.float64)
    to = _buffer(bufs, "type", (Nv,), np.int32)
    go = _buffer(bufs, "tag", (Nv,), np.int32)
    oo = _buffer(bufs, "owner", (Nv,), np.int32)
    io = _buffer(bufs, "img", (Nv, 3), np.int8)
    row_off = np.empty(K + 1, np.int32)
    ghost_off = np.empty(K + 1, np.int32)
    shift = np.empty((K, 3), np.float64)
    _lib.check(L.swr_batch_pack(*args, Nv, xo.ctypes.data, to.ctypes.data, go.ctypes.data, oo.ctypes.data,
                                io.ctypes.data, row_off.ctypes.data, ghost_off.ctypes.data, shift.ctypes.data,
                                C.byref(N)), "swr_batch_pack")
    return Packed(x=xo, type=to, tag=go, owner=oo, img=io, n=n, row_off=row_off, ghost_off=ghost_off, shift=shift)


class HistStore:
    """A system's QEq warm start (s/t histories, 5 per atom each) kept in device
    memory between evaluations in different batches (swr_hist_save/load)."""

    def __init__(self):
        self.h = C.c_void_p(None)
        self.n = 0

    def __del__(self):
        if self.h.value and _lib is not None:
            _lib.lib().swr_hist_free(self.h)
            self.h = C.c_void_p(None)


class Batch:
    """One device system holding a batch; evaluate() runs nbr + QEq + lists + BO (+ the
    non-bonded pass) once for all."""

    def __init__(self, ffield=DEFAULT_FFIELD, elements=ELEMENTS, tol=1e-6, mixed=False):
        self.elements = list(elements)
        self.sys = _lib.System(ffield, self.elements, tol=tol, qeq_mixed=1 if mixed else 0)
        self.packed = None
        self._bufs = {}

    def load(self, batch: list[systems.System], q=None, hist=None):
        """New atoms: list capacities are estimated afresh and QEq starts from
        zero history, except for systems k with a HistStore hist[k] (their
        solve starts from its extrapolation)."""
        P = pack(batch, self.elements, bufs=self._bufs)
        h = self.sys
        h.set_atoms(P.n, P.x, P.type, P.tag, owner=P.owner, img=P.img, q=q)
        _lib.check(h.L.swr_new_atoms(h.h), "swr_new_atoms")
        off = np.ascontiguousarray(P.row_off, dtype=np.int32)
        _lib.check(h.L.swr_set_batch(h.h, len(batch), off.ctypes.data_as(C.POINTER(C.c_int))), "swr_set_batch")
        for k, st in enumerate(hist or []):
            if st is not None:
                assert st.n == batch[k].n, "warm start of another system"
                _lib.check(h.L.swr_hist_load(h.h, int(P.row_off[k]), st.n, st.h), "swr_hist_load")
        self.packed = P
        return P

    def save_history(self, k, store=None, fill=False):
        """Copies system k's s/t histories after the solve into a HistStore
        (reused when given); fill sets all five slots to the solution (a
        first evaluation, whose history holds it in slot 0 only).  Enqueues
        only: sync() before another batch loads it."""
        store = store or HistStore()
        P = self.packed
        n = int(P.row_off[k + 1] - P.row_off[k])
        _lib.check(self.sys.L.swr_hist_save(self.sys.h, int(P.row_off[k]), n, 1 if fill else 0, C.byref(store.h)),
                   "swr_hist_save")
        store.n = n
        return store

    def evaluate(self, nonbonded=False, bonded=False):
        """Neighbour list, QEq (per-system), far/bond/hbond lists, BO; with
        nonbonded the vdW + Coulomb energies and gradient, with bonded the
        bond, atom, valence-angle, torsion and hydrogen-bond energies and the
        bond-order chain rule of the gradient;
        returns max CG iterations (s, t)."""
        h = self.sys
        _lib.check(h.L.swr_npair_full_bin_ghost_build(h.h, C.c_double(CUTGHOST)), "swr_npair_full_bin_ghost_build")
        a, b = C.c_int(0), C.c_int(0)
        _lib.check(h.L.swr_qeq_pre_force(h.h, C.byref(a), C.byref(b)), "swr_qeq_pre_force")
        _lib.check(h.L.swr_pair_compute(h.h), "swr_pair_compute")
        if nonbonded:
            _lib.check(h.L.swr_vdw_coulomb_energy(h.h), "swr_vdw_coulomb_energy")
        if bonded:
            _lib.check(h.L.swr_atom_energy(h.h), "swr_atom_energy")
            _lib.check(h.L.swr_valence_angles(h.h), "swr_valence_angles")
            _lib.check(h.L.swr_torsion_angles(h.h), "swr_torsion_angles")
            _lib.check(h.L.swr_hydrogen_bonds(h.h), "swr_hydrogen_bonds")
            _lib.check(h.L.swr_add_dbond_forces(h.h), "swr_add_dbond_forces")
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

    def energies(self):
        """Per system, its non-bonded (vdW + Coulomb) energy: the sum of its
        atoms' shares (after evaluate(nonbonded=True))."""
        e = self.sys.get("nonbonded_eatom")
        P = self.packed
        return np.array([e[P.row_off[k]:P.row_off[k + 1]].sum() for k in range(len(P.row_off) - 1)])

    def bonded_energies(self):
        """Per system, (e_bond, e_lp, e_ov, e_un, e_pol, e_hb, e_ang, e_pen,
        e_coa, e_tor, e_con) summed over its local atoms (after
        evaluate(bonded=True))."""
        h = self.sys
        stride, n = h.scalar("bonded_stride"), h.scalar("n")
        e = np.vstack([h.get("bonded_eatom").reshape(5, stride)[:, :n], h.get("hbond_eatom")[None, :n],
                       h.get("valence_eatom").reshape(3, -1)[:, :n], h.get("torsion_eatom").reshape(2, -1)[:, :n]])
        P = self.packed
        return np.array([e[:, P.row_off[k]:P.row_off[k + 1]].sum(axis=1) for k in range(len(P.row_off) - 1)])

    def gradients(self):
        """Per system, the energy gradient of its local atoms (n_k x 3) from
        the passes evaluate() ran, ghost entries summed onto their owners (as
        LAMMPS' reverse communication does)."""
        h = self.sys
        P = self.packed
        N = h.scalar("N")
        g = h.get("fCdDelta").reshape(-1, 4)[:N, :3]
        tot = g[:P.n].copy()
        if N > P.n:
            np.add.at(tot, h.get("owner")[P.n:N], g[P.n:N])
        return [tot[P.row_off[k]:P.row_off[k + 1]] for k in range(len(P.row_off) - 1)]

    def iterations(self):
        it = self.sys.get("sys_iters")
        return it.reshape(-1, 2)

    def histories(self):
        """Per system, warm-start s/t histories (n_k x 5 each) for a
        follow-on evaluation of the same atoms: every slot holds the solution
        this solve reached, so fix qeq/reax's extrapolation (cubic in s,
        quadratic in t; fix_qeq_reax_sunway.cpp:711-714) returns it exactly.
        (The raw history after one solve from zero holds it in slot 0 only,
        which those formulas would extrapolate to 4 s and 3 t.)"""
        P = self.packed
        sh = self.sys.get("s_hist").reshape(-1, 5)[:, :1]
        th = self.sys.get("t_hist").reshape(-1, 5)[:, :1]
        return [(np.repeat(sh[P.row_off[k]:P.row_off[k + 1]], 5, axis=1),
                 np.repeat(th[P.row_off[k]:P.row_off[k + 1]], 5, axis=1))
                for k in range(len(P.row_off) - 1)]
