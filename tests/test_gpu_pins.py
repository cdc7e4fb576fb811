"""GPU (HIP, gfx950) checks of the list bookkeeping, the QEq warm start and
the CG recurrence against what the mathematics or a hand count fixes (the
same known answers that pin the oracle), plus the atom-only neighbour list
against the oracle's brute-force rows.  Every call goes through the C-ABI.
"""
import ctypes as C

import numpy as np
import pytest

from sw_reaxff_amd import systems

from .parity import atoms_of, harness, verlet_of, water_c0
from .test_oracle_pins_solver_lists import _crowded_carbon, _dodecahedron, _h_chain, _textbook_pcg

pytestmark = pytest.mark.gpu


def _system(gpu_lib, ffield, elements, x, types, n=None, owner=None, **ctl):
    x = np.asarray(x, dtype=np.float64)
    N = x.shape[0]
    n = N if n is None else n
    owner = np.arange(N, dtype=np.int32) if owner is None else np.asarray(owner, np.int32)
    tags = np.arange(1, N + 1, dtype=np.int32)
    tags[n:] = tags[owner[n:]]
    h = gpu_lib.System(ffield, elements, **ctl)
    h.set_atoms(n, x, np.asarray(types, np.int32), tags, owner=owner)
    h.build_neighbors(12.0)
    return h


def _pair_stages(h):
    for st in ("write_reax_atoms_and_pack", "reset", "write_reax_lists_init_forces", "validate_lists",
               "sort_bonds"):
        h.call(st)
    h.sync()


def test_gpu_warm_start_extrapolation_exact(gpu_lib, ffield):
    """init_matvec's s = 4(s0+s2) - (6 s1 + s3), t = t2 + 3(t0 - t1) returns a
    cubic s history's and a quadratic t history's value at the present step
    (fix_qeq_reax_sunway.cpp:711-714)."""
    k = systems.cluster(60, seed=11)
    h = _system(gpu_lib, ffield, k.elements, k.x, k.type)
    n = k.n
    rng = np.random.default_rng(5)
    cs, ct = rng.normal(size=(n, 4)), rng.normal(size=(n, 3))
    back = -np.arange(1, 6, dtype=np.float64)
    sh = sum(cs[:, [p]] * back[None, :] ** p for p in range(4))
    th = sum(ct[:, [p]] * back[None, :] ** p for p in range(3))
    h.put("s_hist", np.ascontiguousarray(sh))
    h.put("t_hist", np.ascontiguousarray(th))
    h.call("qeq_init_matvec")
    h.sync()
    np.testing.assert_allclose(h.get("s"), cs[:, 0], rtol=0, atol=1e-12 * np.abs(sh).max())
    np.testing.assert_allclose(h.get("t"), ct[:, 0], rtol=0, atol=1e-12 * np.abs(th).max())


def test_gpu_cg_iterates_equal_textbook_pcg(gpu_lib, ffield):
    """The device CG_v2 (s and t together) after imax - 1 updates equals the
    textbook Jacobi PCG's iterate on the same matrix, for the first twelve
    iterations, and stops within one iteration of the textbook's."""
    k = systems.cluster(90, seed=12)
    tol = 1e-10
    n = k.n
    rng = np.random.default_rng(8)
    x0s, x0t = rng.normal(scale=0.1, size=n), rng.normal(scale=0.1, size=n)

    def solve(imax):
        h = _system(gpu_lib, ffield, k.elements, k.x, k.type, tol=tol, qeq_imax=imax)
        # all five history slots equal to x0: the extrapolations return x0
        h.put("s_hist", np.repeat(x0s[:, None], 5, axis=1))
        h.put("t_hist", np.repeat(x0t[:, None], 5, axis=1))
        h.call("qeq_init_matvec")
        its = h.qeq_cg()
        return h, its

    h, _ = solve(1)
    f, nn, j, v = h.get("H_firstnbr"), h.get("H_numnbrs"), h.get("H_jlist"), h.get("H_val")
    A = np.zeros((n, n))
    for i in range(n):
        np.add.at(A[i], j[f[i]:f[i] + nn[i]], v[f[i]:f[i] + nn[i]])
    Minv = h.get("Hdia_inv")
    A += np.diag(1.0 / Minv)
    ref = {}
    for name, b, x0 in (("s", h.get("b_s"), x0s), ("t", h.get("b_t"), x0t)):
        xs, norms = _textbook_pcg(A, b, Minv, x0, 80)
        stop = next(m for m, r in enumerate(norms) if r / np.sqrt(b @ b) <= tol)
        ref[name] = (xs, stop, np.abs(xs[-1]).max(), np.linalg.solve(A, b))
    for imax in (2, 3, 5, 8, 12):
        hk, its = solve(imax)
        assert its == (imax, imax)
        for name in ("s", "t"):
            xs, stop, scale, _ = ref[name]
            np.testing.assert_allclose(hk.get(name), xs[imax - 1], rtol=0, atol=1e-9 * scale)
    hk, its = solve(200)
    for w, name in enumerate(("s", "t")):
        xs, stop, scale, sol = ref[name]
        assert its[w] - 1 in (stop, stop + 1), (name, its, stop)
        np.testing.assert_allclose(hk.get(name), sol, rtol=0, atol=1e-8 * scale)


def test_gpu_bond_capacity_rule_and_reset_offsets(gpu_lib, ffield):
    x, types = _crowded_carbon()
    h = _system(gpu_lib, ffield, ["H", "O", "C"], x, types)
    _pair_stages(h)
    nb = h.get("num_bonds")
    assert (h.get("bond_end") - h.get("bond_index")).tolist() == [14] + [1] * 14
    assert nb.tolist() == [28] + [25] * 14                       # Validate_Lists: max(2 cnt, 25)
    # the slot offsets this step used came from Estimate_Storages' capacities
    assert h.get("bond_index").tolist() == [0, 28, 53, 78, 103, 128, 153, 178, 203, 228, 253, 278, 303,
                                            328, 353]
    assert h.get("Hindex")[:15].tolist() == [-1] * 13 + [0, 1]
    assert h.get("hbond_index").tolist() == [0, 25]
    assert (h.get("hbond_end") - h.get("hbond_index")).tolist() == [12, 12]


@pytest.mark.parametrize("nacc,cap", [(18, 25), (19, 26), (20, 28)])
def test_gpu_hbond_capacity_rule(gpu_lib, ffield, nacc, cap):
    c = np.array([30.0, 30.0, 30.0])
    x = np.vstack([c, c + 6.0 * _dodecahedron()[:nacc]])
    h = _system(gpu_lib, ffield, ["H", "O"], x, np.array([1] + [2] * nacc, np.int32))
    _pair_stages(h)
    assert (h.get("hbond_end") - h.get("hbond_index")).tolist() == [nacc]
    assert h.get("num_hbonds")[0] == cap                         # (int)max(1.4 cnt, 25)


@pytest.mark.parametrize("order,marks", [
    ([1, 2, 3, 4, 5], [0, 1, 2, 3, 4, 5]),
    ([5, 4, 3, 2, 1], [0, 1000, 1000, 1000, 1000, 1]),
    ([2, 1, 4, 3, 5], [0, 1000, 1, 1000, 1000, 1000]),
])
def test_gpu_sort_bonds_single_pass(gpu_lib, ffield, order, marks):
    """sort_bonds_C's one in-order pass (reaxc_forces_sw64.c:932-971), not its
    fixed point, on ghost chains."""
    x = _h_chain(order)
    h = _system(gpu_lib, ffield, ["H"], x, np.ones(6, np.int32), n=1, owner=np.zeros(6, np.int32))
    _pair_stages(h)
    assert h.get("bond_index").tolist() == [0, 25, 50, 75, 100, 125]
    assert h.get("bond_mark").tolist() == marks


def test_gpu_atomonly_list_equals_oracle_local_rows(gpu_lib, oracle_lib, ffield):
    """npair full/bin/atomonly (npair_full_bin_atomonly_sunway.cpp:39-203,
    npair_full_bin_atomonly_sw5.c:57-180) against the oracle's brute-force
    enumeration: every local row equals the oracle's row as a set, ghost rows
    are empty (gnum = 0), j != i and every pair within the cutoff."""
    s = water_c0(seed=4)
    lmp = harness(s, ffield)
    h = lmp.system
    A = atoms_of(h)
    n, N = A["n"], A["N"]
    gpu_lib.check(h.L.swr_npair_full_bin_atomonly_build(h.h, C.c_double(12.0)), "atomonly")
    h.sync()
    aoff, alst = verlet_of(h)
    assert aoff.size == N + 1 and np.all(np.diff(aoff)[n:] == 0)
    o = oracle_lib.Oracle(ffield, s.elements, threads=8)
    o.set_atoms(n, A["x"], A["type"], A["tag"], A["owner"])
    o.build_verlet(12.0, brute=True)
    ooff, olst = o.get("verlet_offsets"), o.get("verlet_list")
    assert np.array_equal(np.diff(aoff)[:n], np.diff(ooff)[:n])
    gi = np.repeat(np.arange(n), np.diff(aoff)[:n])
    oi = np.repeat(np.arange(n), np.diff(ooff)[:n])
    g = np.lexsort((alst[:aoff[n]], gi))
    w = np.lexsort((olst[:ooff[n]], oi))
    assert np.array_equal(alst[:aoff[n]][g], olst[:ooff[n]][w])
    assert not np.any(alst[:aoff[n]] == gi)


def test_gpu_verlet_bins_beyond_one_wave(gpu_lib, oracle_lib, ffield):
    """The Verlet build takes a bin's atoms 64 at a time (one per lane) and
    streams the bin's stencil past them: a dense random cloud of 1,500 atoms
    in a 20-A cube at a 16-A cutoff (two 10-A bins a side, ~190 atoms each,
    so three groups per bin) against the oracle's brute-force enumeration
    (npair_full_bin_ghost_sw5.c:183-199): every row equal as a set, j != i."""
    rng = np.random.default_rng(7)
    n = 1500
    x = rng.uniform(0.0, 20.0, size=(n, 3))
    types = np.ones(n, np.int32)   # LAMMPS types, 1-based
    h = _system(gpu_lib, ffield, ["O"], x, types)
    h.build_neighbors(16.0)
    h.sync()
    off, lst = verlet_of(h)
    o = oracle_lib.Oracle(ffield, ["O"], threads=8)
    o.set_atoms(n, x, types, np.arange(1, n + 1, dtype=np.int32), np.arange(n, dtype=np.int32))
    o.build_verlet(16.0, brute=True)
    ooff, olst = o.get("verlet_offsets"), o.get("verlet_list")
    cnt = np.diff(off)
    print(f"[verlet] {n} atoms, {off[-1]} pairs, rows {cnt.min()}..{cnt.max()}", flush=True)
    assert np.array_equal(cnt, np.diff(ooff))
    rows = np.repeat(np.arange(n), cnt)
    assert not np.any(lst[:off[-1]] == rows)
    g = np.lexsort((lst[:off[-1]], rows))
    w = np.lexsort((olst[:ooff[-1]], rows))
    assert np.array_equal(lst[:off[-1]][g], olst[:ooff[-1]][w])
