"""numpy restatement of the batch packing that swr_batch_pack does in C++:
ghost images in LAMMPS Comm::borders order (systems.borders), uniform cubic
cells of edge (largest extent + cutghost + 1), local atoms of all systems
first, then their ghosts.  Used by tests/test_host.py as the reference."""
import numpy as np

from sw_reaxff_amd import systems
from sw_reaxff_amd.batch import Packed


def ghosts_of(s: systems.System, cutghost):
    """Local + ghost coordinates, owner and periodic image of one system."""
    if not s.periodic:
        n = s.n
        return s.x.copy(), np.arange(n, dtype=np.int32), np.zeros((n, 3), np.int8)
    xa, owner = systems.borders(s.x, s.lo, s.hi, cutghost)
    L = s.hi - s.lo
    img = np.rint((xa - s.x[owner]) / L).astype(np.int8)
    return xa, owner.astype(np.int32), img


def pack(batch: list[systems.System], elements, cutghost) -> Packed:
    """numpy restatement of swr_batch_pack (the test reference)."""
    K = len(batch)
    parts = [ghosts_of(s, cutghost) for s in batch]
    ext = []
    for s, (xa, _, _) in zip(batch, parts):
        lo = np.minimum(xa.min(0), s.lo) if s.n else s.lo
        hi = np.maximum(xa.max(0), s.hi) if s.n else s.hi
        ext.append((lo, hi))
    # uniform cubic cells on an m x m x m grid, each a system's extent plus the cutoff gap
    size = max(float(np.max(hi - lo)) for lo, hi in ext) + cutghost + 1.0
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


