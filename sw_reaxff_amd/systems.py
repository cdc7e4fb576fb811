"""Seeded synthetic inputs for the nbr + BO + QEq path.

Builders for the periodic rigid-geometry water boxes and the non-periodic
water/methanol clusters the benchmark and tests run on.  Every builder is a
pure function of its arguments and seed.  Units follow LAMMPS `units real`:
positions in Angstrom, velocities in Angstrom/fs, masses in g/mol.

These produce *inputs* only (what a LAMMPS data file or create_atoms would
provide); nothing here is part of the device path.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

AVOGADRO = 6.02214076e23
KB_J = 1.380649e-23          # J/K
AMU_KG = 1.66053906660e-27   # kg
MASS = {"H": 1.008, "C": 12.011, "O": 15.999, "N": 14.007}
WATER_MASS = 2 * MASS["H"] + MASS["O"]
METHANOL_MASS = MASS["C"] + MASS["O"] + 4 * MASS["H"]

# Rigid water geometry: O-H 0.9572 A, H-O-H 104.52 deg.
WATER_OH = 0.9572
WATER_HOH = math.radians(104.52)
# Methanol geometry: C-O 1.43, O-H 0.96, C-H 1.09 A, tetrahedral angles.
MEOH_CO, MEOH_OH, MEOH_CH = 1.43, 0.96, 1.09
TETRA = math.acos(-1.0 / 3.0)


@dataclass
class System:
    """A set of atoms in LAMMPS conventions (types are 1-based)."""
    x: np.ndarray            # [n,3] float64
    v: np.ndarray            # [n,3] float64, A/fs
    type: np.ndarray         # [n] int32, LAMMPS type 1..len(elements)
    tag: np.ndarray          # [n] int32, 1-based atom IDs
    elements: list           # element name of each LAMMPS type
    lo: np.ndarray           # [3]
    hi: np.ndarray           # [3]
    periodic: bool
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    @property
    def masses(self) -> np.ndarray:
        return np.array([MASS[e] for e in self.elements])[self.type - 1]


def _random_rotations(rng: np.random.Generator, m: int) -> np.ndarray:
    """Uniform random rotation matrices from normalised Gaussian quaternions."""
    qv = rng.normal(size=(m, 4))
    qv /= np.linalg.norm(qv, axis=1, keepdims=True)
    w, x, y, z = qv.T
    R = np.empty((m, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def water_body() -> np.ndarray:
    """Body-frame water: O at the origin, H in the xy plane."""
    h = WATER_HOH / 2
    return np.array([[0.0, 0.0, 0.0],
                     [WATER_OH * math.sin(h), WATER_OH * math.cos(h), 0.0],
                     [-WATER_OH * math.sin(h), WATER_OH * math.cos(h), 0.0]])


def methanol_body() -> np.ndarray:
    """Body-frame methanol: C, O, H(O), 3 H(C); tetrahedral angles at C and O."""
    c = np.zeros(3)
    o = np.array([0.0, 0.0, MEOH_CO])
    # hydroxyl H on O, angle C-O-H tetrahedral, in the xz plane
    ho = o + MEOH_OH * np.array([math.sin(math.pi - TETRA), 0.0, -math.cos(math.pi - TETRA)])
    hs = []
    for k in range(3):
        phi = 2 * math.pi * k / 3 + math.pi
        hs.append(MEOH_CH * np.array([math.sin(TETRA) * math.cos(phi), math.sin(TETRA) * math.sin(phi),
                                      math.cos(TETRA)]))
    return np.vstack([c, o, ho] + hs)


def box_length(nmol: int, mol_mass: float, density_g_cm3: float) -> float:
    """Cube edge (A) holding nmol molecules of mol_mass at the given density."""
    vol_cm3 = nmol * mol_mass / AVOGADRO / density_g_cm3
    return (vol_cm3 * 1e24) ** (1.0 / 3.0)


def maxwell_velocities(rng: np.random.Generator, masses: np.ndarray, temp: float) -> np.ndarray:
    """Maxwell-Boltzmann velocities (A/fs) at temp, centre-of-mass momentum removed."""
    sigma = np.sqrt(KB_J * temp / (masses * AMU_KG)) * 1e-5
    v = rng.normal(size=(masses.size, 3)) * sigma[:, None]
    vcm = (v * masses[:, None]).sum(0) / masses.sum()
    return v - vcm


def water_box(nmol: int, seed: int, density: float = 1.0, temp: float = 300.0) -> System:
    """Periodic cubic box of rigid waters on a simple-cubic lattice.

    The lattice has m = ceil(nmol^(1/3)) sites per edge at spacing L/m; nmol
    sites are chosen by a seeded permutation (kept in ascending site order),
    each molecule gets a seeded uniform random orientation with O on its site,
    and Maxwell velocities at temp are drawn from the same seed.  Atom order
    per molecule is O, H, H.  LAMMPS types: 1 = H, 2 = O.
    """
    rng = np.random.default_rng(seed)
    L = box_length(nmol, WATER_MASS, density)
    m = int(math.ceil(round(nmol ** (1.0 / 3.0), 9)))
    while m ** 3 < nmol:
        m += 1
    a = L / m
    sites = np.sort(rng.permutation(m ** 3)[:nmol])
    iz, rem = np.divmod(sites, m * m)
    iy, ix = np.divmod(rem, m)
    centers = (np.stack([ix, iy, iz], axis=1) + 0.5) * a
    R = _random_rotations(rng, nmol)
    body = water_body()
    pos = centers[:, None, :] + np.einsum("mij,kj->mki", R, body)
    x = np.mod(pos.reshape(-1, 3), L)
    types = np.tile(np.array([2, 1, 1], dtype=np.int32), nmol)
    elements = ["H", "O"]
    masses = np.array([MASS[e] for e in elements])[types - 1]
    v = maxwell_velocities(rng, masses, temp)
    tag = np.arange(1, 3 * nmol + 1, dtype=np.int32)
    return System(x=x, v=v, type=types, tag=tag, elements=elements,
                  lo=np.zeros(3), hi=np.full(3, L), periodic=True,
                  meta={"kind": "water_box", "nmol": nmol, "seed": seed, "L": L,
                        "lattice": m, "density": density, "temp": temp})


def cluster(natoms: int, seed: int, density: float = 0.9, min_heavy: float = 2.2,
            pad: float = 12.0, temp: float = 300.0) -> System:
    """Non-periodic water/methanol cluster of about natoms atoms.

    Molecules (water or methanol, 50/50 by seeded draw) are placed by seeded
    random sequential insertion inside a sphere sized for the target density,
    rejecting placements whose heavy atoms (O, C) come closer than min_heavy
    to an already placed heavy atom.  The bounding box is padded by pad A.
    LAMMPS types: 1 = H, 2 = O, 3 = C.
    """
    rng = np.random.default_rng(seed)
    wb, mb = water_body(), methanol_body()
    kinds = []
    count = 0
    while count < natoms:
        k = 0 if rng.random() < 0.5 else 1
        kinds.append(k)
        count += 3 if k == 0 else 6
    mass_tot = sum(WATER_MASS if k == 0 else METHANOL_MASS for k in kinds)
    vol = mass_tot / AVOGADRO / density * 1e24
    radius = (3 * vol / (4 * math.pi)) ** (1.0 / 3.0)
    # placed heavy atoms in a hash of cubic cells of edge min_heavy: every
    # heavy atom closer than min_heavy to a candidate lies in the 27 cells
    # around it, so the test below rejects exactly the placements a check
    # against all placed heavy atoms would
    cells: dict = {}
    inv = 1.0 / min_heavy
    lim = min_heavy * min_heavy
    xs, ts = [], []
    for k in kinds:
        body = wb if k == 0 else mb
        types = np.array([2, 1, 1]) if k == 0 else np.array([3, 2, 1, 1, 1, 1])
        hv = types >= 2
        placed = False
        r_try = radius
        for attempt in range(2000):
            if attempt and attempt % 200 == 0:
                r_try *= 1.05
            c = rng.uniform(-1, 1, size=3)
            if np.dot(c, c) > 1:
                continue
            c *= r_try
            Rm = _random_rotations(rng, 1)[0]
            p = c + body @ Rm.T
            ph = p[hv]
            keys = [tuple(v) for v in np.floor(ph * inv).astype(np.int64).tolist()]
            clash = False
            for a, (ix, iy, iz) in zip(ph, keys):
                for dx in (-1, 0, 1):
                    for dy in (-1, 0, 1):
                        for dz in (-1, 0, 1):
                            near = cells.get((ix + dx, iy + dy, iz + dz))
                            if near is not None and (((np.asarray(near) - a) ** 2).sum(-1) < lim).any():
                                clash = True
                                break
                        if clash:
                            break
                    if clash:
                        break
                if clash:
                    break
            if clash:
                continue
            for a, key in zip(ph, keys):
                cells.setdefault(key, []).append(a)
            xs.append(p)
            ts.append(types)
            placed = True
            break
        if not placed:
            raise RuntimeError("cluster packing failed")
    x = np.vstack(xs)
    types = np.concatenate(ts).astype(np.int32)
    lo = x.min(0) - pad
    hi = x.max(0) + pad
    elements = ["H", "O", "C"]
    masses = np.array([MASS[e] for e in elements])[types - 1]
    v = maxwell_velocities(rng, masses, temp)
    tag = np.arange(1, x.shape[0] + 1, dtype=np.int32)
    return System(x=x, v=v, type=types, tag=tag, elements=elements, lo=lo, hi=hi,
                  periodic=False, meta={"kind": "cluster", "natoms": int(x.shape[0]),
                                        "seed": seed, "molecules": len(kinds)})


def cluster_sizes(count: int, lo: int, hi: int, seed: int) -> np.ndarray:
    """Log-uniform integer sizes on [lo, hi]."""
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(math.log(lo), math.log(hi), size=count)).astype(np.int64)


def borders(x: np.ndarray, lo: np.ndarray, hi: np.ndarray, cutghost: float):
    """Ghost images of a periodic orthogonal box, in LAMMPS Comm::borders order.

    For each dimension, two swaps over the atoms present at the start of that
    dimension: atoms with x <= lo+cutghost are imaged at +L, then atoms with
    x >= hi-cutghost are imaged at -L.  Returns (x_all, owner) where owner[g]
    is the local atom ghost g copies (owner[i] = i for local atoms).
    """
    L = hi - lo
    xa = x.copy()
    owner = np.arange(x.shape[0], dtype=np.int32)
    for d in range(3):
        nlast = xa.shape[0]
        sel0 = np.nonzero(xa[:nlast, d] <= lo[d] + cutghost)[0]
        sel1 = np.nonzero(xa[:nlast, d] >= hi[d] - cutghost)[0]
        g0 = xa[sel0].copy()
        g0[:, d] += L[d]
        g1 = xa[sel1].copy()
        g1[:, d] -= L[d]
        xa = np.vstack([xa, g0, g1])
        owner = np.concatenate([owner, owner[sel0], owner[sel1]])
    return xa, owner
