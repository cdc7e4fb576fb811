"""Host-side logic on the CPU: input builders, ghost images, the command
interface and the C-ABI library's exported symbols (no GPU calls)."""
import math
import os
import re
import subprocess

import numpy as np
import pytest

from sw_reaxff_amd import systems

from .conftest import FFIELD

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "sw_reaxff_amd", "libswreax.so")


# ------------------------------------------------------------------ builders
def test_water_box_geometry_and_density():
    s = systems.water_box(648, seed=1)
    assert s.n == 1944 and s.elements == ["H", "O"]
    L = s.hi - s.lo
    assert np.allclose(L, 26.86, atol=0.01)                     # 1.0 g/cm^3
    x = s.x.reshape(-1, 3, 3)                                   # O, H, H per molecule
    assert np.all(s.type.reshape(-1, 3) == [2, 1, 1])
    oh1 = np.linalg.norm(x[:, 1] - x[:, 0], axis=1)
    oh2 = np.linalg.norm(x[:, 2] - x[:, 0], axis=1)
    assert np.allclose(oh1, 0.9572) and np.allclose(oh2, 0.9572)
    c = np.sum((x[:, 1] - x[:, 0]) * (x[:, 2] - x[:, 0]), axis=1) / (oh1 * oh2)
    assert np.allclose(np.degrees(np.arccos(c)), 104.52)
    assert np.all(s.x >= s.lo) and np.all(s.x < s.hi)
    assert np.array_equal(s.tag, np.arange(1, s.n + 1))
    # 300 K Maxwell-Boltzmann, zero total momentum
    p = (s.masses[:, None] * s.v).sum(axis=0)
    assert np.allclose(p, 0.0, atol=1e-10)


def test_builders_are_seeded():
    a, b = systems.water_box(200, seed=9), systems.water_box(200, seed=9)
    c = systems.water_box(200, seed=10)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.v, b.v)
    assert not np.array_equal(a.x, c.x)
    k1, k2 = systems.cluster(400, seed=4), systems.cluster(400, seed=4)
    assert np.array_equal(k1.x, k2.x) and np.array_equal(k1.type, k2.type)


def test_cluster_packing():
    k = systems.cluster(900, seed=2)
    assert not k.periodic
    # molecules are consecutive: water starts with O (3 atoms), methanol with C (6)
    mol, i, m = np.empty(k.n, int), 0, 0
    while i < k.n:
        size = 3 if k.type[i] == 2 else 6
        assert k.type[i] in (2, 3)
        mol[i:i + size] = m
        i, m = i + size, m + 1
    assert m == k.meta["molecules"]
    hv = np.isin(np.array(k.elements)[k.type - 1], ["C", "O"])
    heavy, hm = k.x[hv], mol[hv]
    d = np.sqrt(((heavy[:, None, :] - heavy[None, :, :]) ** 2).sum(-1))
    d[hm[:, None] == hm[None, :]] = np.inf
    assert d.min() >= 2.2 - 1e-9
    # padded by 12 A: no atom within a ghost cut of a face
    assert np.all(k.x - k.lo >= 12.0 - 1e-9) and np.all(k.hi - k.x >= 12.0 - 1e-9)
    sizes = systems.cluster_sizes(512, 50, 3000, seed=0)
    assert sizes.min() >= 50 and sizes.max() <= 3000
    # log-uniform on [50, 3000]: geometric mean sqrt(50*3000) = 387, arithmetic
    # mean (3000-50)/ln(60) = 720; bounds are 3 standard errors for 512 draws
    gm = np.exp(np.log(sizes).mean())
    assert 387 * math.exp(-0.157) < gm < 387 * math.exp(0.157)
    assert 620 < sizes.mean() < 820


# ------------------------------------------------------------------ ghosts
def test_borders_equals_brute_force_images():
    """Every periodic image inside [lo - cut, hi + cut) is a ghost, exactly once
    (LAMMPS Comm::borders on one domain)."""
    s = systems.water_box(300, seed=5)
    cut = 6.0
    xa, owner = systems.borders(s.x, s.lo, s.hi, cut)
    n = s.n
    L = s.hi - s.lo
    assert np.array_equal(owner[:n], np.arange(n))
    shifts = np.rint((xa - s.x[owner]) / L).astype(int)
    assert np.allclose(xa, s.x[owner] + shifts * L)
    got = set(zip(owner[n:].tolist(), map(tuple, shifts[n:].tolist())))
    assert len(got) == xa.shape[0] - n                            # no duplicates
    want = set()
    for sx in (-1, 0, 1):
        for sy in (-1, 0, 1):
            for sz in (-1, 0, 1):
                if sx == sy == sz == 0:
                    continue
                y = s.x + np.array([sx, sy, sz]) * L
                ok = np.all((y >= s.lo - cut) & (y <= s.hi + cut), axis=1)
                # the borders test is on the coordinate before the shift
                sel = np.ones(n, bool)
                for d, sd in enumerate((sx, sy, sz)):
                    if sd == 1:
                        sel &= s.x[:, d] <= s.lo[d] + cut
                    elif sd == -1:
                        sel &= s.x[:, d] >= s.hi[d] - cut
                want |= {(int(i), (sx, sy, sz)) for i in np.nonzero(sel & ok)[0]}
    assert got == want


def test_borders_order_is_per_dimension_swaps():
    s = systems.water_box(200, seed=6)
    xa, owner = systems.borders(s.x, s.lo, s.hi, 5.0)
    n = s.n
    first = xa[n:]
    # the first ghosts are the x-swap of local atoms near lo (shift +L_x), in index order
    sel0 = np.nonzero(s.x[:, 0] <= s.lo[0] + 5.0)[0]
    assert np.array_equal(owner[n:n + sel0.size], sel0)
    assert np.allclose(first[:sel0.size, 0], s.x[sel0, 0] + (s.hi - s.lo)[0])


# ------------------------------------------------------------------ C-ABI
def _declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", fn)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        txt = re.sub(r"//[^\n]*", "", txt)
        for m in re.finditer(r"\b(swr_[a-z0-9_]+)\s*\(", txt):
            names.add(m.group(1))
    return names


@pytest.fixture(scope="module")
def built_lib():
    if not os.path.exists(LIB):
        subprocess.run(["make", "-C", os.path.join(ROOT, "sw_reaxff_amd", "csrc"), "-j8"], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB


def test_library_exports_every_declared_symbol(built_lib):
    import ctypes
    lib = ctypes.CDLL(built_lib)
    declared = _declared_symbols()
    assert len(declared) >= 30
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    from sw_reaxff_amd import _lib
    assert set(_lib.EXPORTS) <= declared


# ------------------------------------------------------------------ command interface
@pytest.fixture()
def lmp(built_lib):
    from sw_reaxff_amd.lammps import Lammps
    return Lammps()


def test_standard_script_parses(lmp):
    from sw_reaxff_amd.lammps import standard_script
    lmp.commands(standard_script(["H", "O"]))


@pytest.mark.parametrize("cmd,msg", [
    ("pair_style lj/cut 10", "Unknown pair style"),
    ("pair_style reax/c NULL checkqeq maybe", "Illegal pair_style"),
    ("pair_style reax/c NULL lgvdw yes", "lgvdw"),
    ("pair_style reax/c NULL safezone -1", "safezone"),
    ("pair_coeff 1 1 ffield.reax H", "Incorrect args"),
    ("fix 1 all qeq/reax 1 0.0 10.0 1e-6", "Illegal fix qeq/reax"),
    ("fix 1 all qeq/reax 0 0.0 10.0 1e-6 reax/c", "Illegal fix qeq/reax"),
    ("fix 1 all qeq/reax 1 0.0 10.0 1e-6 reax/c dual", "Dual keyword"),
    ("fix 1 all qeq/reax 1 0.0 -1.0 1e-6 reax/c", "negative upper Taper"),
    ("neighbor 2.0 nsq", "Illegal neighbor"),
    ("units metal", "units real"),
    ("run_style respa", "Unknown command"),
])
def test_command_errors(lmp, cmd, msg):
    from sw_reaxff_amd._lib import SwrError
    with pytest.raises(SwrError) as ei:
        lmp.command(cmd)
    assert ei.value.code == 1
    assert msg in str(ei.value)


@pytest.mark.parametrize("pair,fix", [
    ("reax/c/sunway", "qeq/reax/sunway"),      # PairStyle pair_reaxc_sunway.h:26, FixStyle fix_qeq_reax_sunway.h:26
    ("reax/c", "qeq/reax"),                    # the names a -sf sunway input uses
])
def test_sunway_style_names(built_lib, pair, fix):
    """The reference's command surface: its style names and the suffix
    command are accepted with the same arguments and errors."""
    from sw_reaxff_amd._lib import SwrError
    from sw_reaxff_amd.lammps import Lammps, standard_script
    L = Lammps()
    L.command("suffix sunway")
    L.command(f"pair_style {pair} NULL checkqeq yes safezone 1.6 mincap 100")
    L.command(f"pair_coeff * * {FFIELD} H O")
    L.command(f"fix 2 all {fix} 1 0.0 10.0 1e-6 reax/c")
    L.command("fix 1 all nve/sunway")
    with pytest.raises(SwrError) as ei:
        L.command(f"fix 3 all {fix} 1 0.0 10.0 1e-6 reax/c dual")
    assert "Dual keyword" in str(ei.value)
    with pytest.raises(SwrError) as ei:
        L.command(f"pair_style {pair}/omp NULL")
    assert "Unknown pair style" in str(ei.value)
    script = standard_script(["H", "O"])
    assert "pair_style reax/c/sunway NULL" in script and "qeq/reax/sunway" in script


def test_setup_without_atoms_is_an_error(lmp):
    from sw_reaxff_amd._lib import SwrError
    from sw_reaxff_amd.lammps import standard_script
    lmp.commands(standard_script(["H", "O"]))
    with pytest.raises(SwrError):
        lmp.setup()


# ------------------------------------------------------------------ batch packing (host C++)
def _pack_mix():
    return [systems.cluster(60, seed=5), systems.water_box(64, seed=6),      # L ~ 12.4 A < 2 x cutghost
            systems.water_box(300, seed=7), systems.cluster(400, seed=8), systems.water_box(100, seed=9)]


@pytest.mark.parametrize("cut", [10.0, 12.0])
def test_batch_pack_equals_numpy_restatement(built_lib, cut):
    from sw_reaxff_amd import batch as B
    from tests import batch_pack_ref as R
    parts = _pack_mix()
    P1 = R.pack(parts, B.ELEMENTS, cut)
    bufs = {}
    for _ in range(2):                                          # second pass reuses the buffers
        P2 = B.pack(parts, cutghost=cut, bufs=bufs)
        assert P1.n == P2.n
        for f in ("x", "type", "tag", "owner", "img", "row_off", "ghost_off", "shift"):
            assert np.array_equal(getattr(P1, f), getattr(P2, f)), f


def test_batch_pack_separates_systems(built_lib):
    from sw_reaxff_amd import batch as B
    parts = _pack_mix()
    P = B.pack(parts, cutghost=10.0)
    # system id of every packed atom (locals, then ghosts)
    sid = np.empty(P.x.shape[0], np.int64)
    for k in range(len(parts)):
        sid[P.row_off[k]:P.row_off[k + 1]] = k
        sid[P.n + P.ghost_off[k]:P.n + P.ghost_off[k + 1]] = k
    assert np.array_equal(sid[P.owner], sid)                    # ghosts copy their own system's atoms
    for k in range(len(parts)):
        a = P.x[sid == k]
        b = P.x[sid != k]
        d2 = ((a[:, None, :] - b[None, :, :]) ** 2).sum(-1)
        assert d2.min() > 10.0 ** 2, k
    # a ghost sits at its owner plus its periodic image
    g = np.arange(P.n, P.x.shape[0])
    L = np.array([parts[k].hi - parts[k].lo for k in sid[g]])
    assert np.allclose(P.x[g], P.x[P.owner[g]] + P.img[g] * L, atol=1e-9)


def test_batch_pack_rejects_bad_input(built_lib):
    import ctypes as C
    from sw_reaxff_amd import _lib
    L = _lib.lib()
    nat = np.array([3], np.int32)
    x = np.zeros((3, 3))
    t = np.ones(3, np.int32)
    lo = np.zeros(3)
    hi = np.zeros(3)                                             # empty periodic box
    per = np.ones(1, np.int32)
    N = C.c_longlong(0)
    rc = L.swr_batch_pack(1, nat.ctypes.data, x.ctypes.data, t.ctypes.data, t.ctypes.data, lo.ctypes.data,
                          hi.ctypes.data, per.ctypes.data, C.c_double(10.0), 0, None, None, None, None, None,
                          None, None, None, C.byref(N))
    assert rc != 0
    hi = np.full(3, 20.0)
    rc = L.swr_batch_pack(1, nat.ctypes.data, x.ctypes.data, t.ctypes.data, t.ctypes.data, lo.ctypes.data,
                          hi.ctypes.data, per.ctypes.data, C.c_double(10.0), 0, None, None, None, None, None,
                          None, None, None, C.byref(N))
    assert rc == 0 and N.value == 3 * 8                          # the corner atom: 2 x 2 x 2 images
    out = np.empty((N.value, 3))
    rc = L.swr_batch_pack(1, nat.ctypes.data, x.ctypes.data, t.ctypes.data, t.ctypes.data, lo.ctypes.data,
                          hi.ctypes.data, per.ctypes.data, C.c_double(10.0), N.value - 1, out.ctypes.data,
                          None, None, None, None, None, None, None, C.byref(N))
    assert rc != 0                                               # capacity below N
    hi = np.array([20.0, 20.0, 9.5])                             # L_z below the 10 A ghost cut
    rc = L.swr_batch_pack(1, nat.ctypes.data, x.ctypes.data, t.ctypes.data, t.ctypes.data, lo.ctypes.data,
                          hi.ctypes.data, per.ctypes.data, C.c_double(10.0), 0, None, None, None, None, None,
                          None, None, None, C.byref(N))
    assert rc == 1 and "shorter than the ghost cutoff" in L.swr_last_error().decode()
    per0 = np.zeros(1, np.int32)                                 # the same box, not periodic: no images
    rc = L.swr_batch_pack(1, nat.ctypes.data, x.ctypes.data, t.ctypes.data, t.ctypes.data, lo.ctypes.data,
                          hi.ctypes.data, per0.ctypes.data, C.c_double(10.0), 0, None, None, None, None, None,
                          None, None, None, C.byref(N))
    assert rc == 0 and N.value == 3
