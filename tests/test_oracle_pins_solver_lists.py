"""Pins of the oracle's QEq warm start and solver, and of its list
bookkeeping (CPU only), each by what the mathematics or a hand count fixes:

  a12  the warm-start extrapolation (fix_qeq_reax_sunway.cpp:711-714) is
       exact on a cubic s history and a quadratic t history;
  a15  CG_v2 (fix_qeq_reax_sunway.cpp:983-1167) is a pipelined form of
       Jacobi-preconditioned CG: in exact arithmetic its iterates are those
       of the textbook recurrence, so the oracle's x after every iteration
       count, and the iteration at which it stops, must be the numpy
       textbook PCG's on the same system;
  a4   Reset (reaxc_reset_tools_sunway.cpp:36-57, :134-197): Hindex and the
       bond/hbond slot offsets of hand-built systems, written out by hand;
  a7   the capacity rules of Estimate_Storages (reaxc_forces_sunway.cpp:
       1159-1294) and Validate_Lists_C (reaxc_forces_sw64.c:866-930),
       num_bonds = max(2 cnt, 25) and num_hbonds = (int)max(1.4 cnt, 25), on
       geometries whose counts are fixed by construction (checked against
       the closed-form BO' of the raw force-field text); sort_bonds_C's
       single in-order pass (:932-971) on ghost chains where one pass and
       the fixed point differ.
"""
import numpy as np
import pytest

from sw_reaxff_amd import systems

from .ffield_kat import FField


@pytest.fixture(scope="module")
def ff(ffield):
    return FField(ffield)


def _atoms(oracle_lib, ffield, elements, x, types, n=None, owner=None, tol=1e-10):
    x = np.asarray(x, dtype=np.float64)
    N = x.shape[0]
    n = N if n is None else n
    owner = np.arange(N, dtype=np.int32) if owner is None else np.asarray(owner, np.int32)
    o = oracle_lib.Oracle(ffield, elements, tol=tol, threads=1)
    tags = np.arange(1, N + 1, dtype=np.int32)
    tags[n:] = tags[owner[n:]]
    o.set_atoms(n, x, np.asarray(types, np.int32), tags, owner)
    o.build_verlet(12.0)
    return o


def _lists(o, estimate=True):
    """pack, far list, (Estimate_Storages), Reset, far list, Init_Forces."""
    o.stage("write_reax_atoms")
    o.stage("write_reax_lists")
    if estimate:
        o.stage("estimate_storages")
    assert o.stage("reset") == 0
    o.stage("write_reax_lists")
    assert o.stage("init_forces") == 0


# ------------------------------------------------------------------ a12 extrapolation
def test_warm_start_extrapolation_exact_on_polynomials(oracle_lib, ffield):
    """s_hist/t_hist[i][k] hold the solutions k+1 steps back; init_matvec
    extrapolates to the present step.  On s(step) a cubic and t(step) a
    quadratic polynomial the extrapolation must return the polynomial's value
    at step 0."""
    k = systems.cluster(60, seed=11)
    o = _atoms(oracle_lib, ffield, k.elements, k.x, k.type)
    n = k.n
    rng = np.random.default_rng(5)
    cs = rng.normal(size=(n, 4))              # s(u) = c0 + c1 u + c2 u^2 + c3 u^3
    ct = rng.normal(size=(n, 3))              # t(u) = c0 + c1 u + c2 u^2
    back = -np.arange(1, 6, dtype=np.float64)  # history slot k is step -(k+1)
    sh = sum(cs[:, [p]] * back[None, :] ** p for p in range(4))
    th = sum(ct[:, [p]] * back[None, :] ** p for p in range(3))
    o.set("s_hist", np.ascontiguousarray(sh))
    o.set("t_hist", np.ascontiguousarray(th))
    assert o.stage("qeq_init_matvec") == 0
    np.testing.assert_allclose(o.get("s")[:n], cs[:, 0], rtol=0, atol=1e-12 * np.abs(sh).max())
    np.testing.assert_allclose(o.get("t")[:n], ct[:, 0], rtol=0, atol=1e-12 * np.abs(th).max())
    # one order higher is not reproduced: the rule is exactly cubic / quadratic
    sh4 = sh + back[None, :] ** 4
    o.set("s_hist", np.ascontiguousarray(sh4))
    o.set("t_hist", np.ascontiguousarray(th + back[None, :] ** 3))
    o.stage("qeq_init_matvec")
    assert np.all(np.abs(o.get("s")[:n] - cs[:, 0]) > 1.0)
    assert np.all(np.abs(o.get("t")[:n] - ct[:, 0]) > 1.0)


# ------------------------------------------------------------------ a15 CG_v2
def _textbook_pcg(A, b, Minv, x0, steps):
    """Jacobi-preconditioned CG (Saad, Iterative Methods, Alg. 9.1): returns
    the iterates x_0..x_steps and the preconditioned residual norms
    sqrt(r_m . M^-1 r_m)."""
    x = x0.copy()
    r = b - A @ x
    z = Minv * r
    p = z.copy()
    rz = r @ z
    xs, norms = [x.copy()], [np.sqrt(rz)]
    for _ in range(steps):
        Ap = A @ p
        alpha = rz / (p @ Ap)
        x = x + alpha * p
        r = r - alpha * Ap
        z = Minv * r
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
        xs.append(x.copy())
        norms.append(np.sqrt(max(rz, 0.0)))
    return xs, norms


@pytest.mark.parametrize("which", [0, 1])
def test_cg_v2_iterates_equal_textbook_pcg(oracle_lib, ffield, which):
    k = systems.cluster(90, seed=12)
    tol = 1e-10
    o = _atoms(oracle_lib, ffield, k.elements, k.x, k.type, tol=tol)
    n = k.n
    assert o.stage("qeq_init_matvec") == 0
    f, nn, j, v = o.get("H_firstnbr"), o.get("H_numnbrs"), o.get("H_jlist"), o.get("H_val")
    A = np.zeros((n, n))
    for i in range(n):
        np.add.at(A[i], j[f[i]:f[i] + nn[i]], v[f[i]:f[i] + nn[i]])
    Minv = o.get("Hdia_inv")[:n]
    A += np.diag(1.0 / Minv)
    b = o.get("b_s" if which == 0 else "b_t")[:n]
    name = "s" if which == 0 else "t"
    rng = np.random.default_rng(7 + which)
    x0 = rng.normal(scale=0.1, size=n)        # a warm start that is not the answer
    xs, norms = _textbook_pcg(A, b, Minv, x0, 60)
    bnorm = np.sqrt(b @ b)
    stop = next(m for m, r in enumerate(norms) if r / bnorm <= tol)
    assert 5 < stop < 60
    scale = np.abs(xs[-1]).max()
    # The first iterates: the two recurrences differ only by rounding (a
    # slipped alpha, beta or dot product moves them at O(step) from the
    # second iteration on).  Later iterates of any two finite-precision CG
    # forms drift apart (measured 1e-5 of the step near iteration 20 on this
    # system) and meet again at convergence.
    for imax in range(1, 13):
        o.set(name, np.ascontiguousarray(x0))
        o.set_cg_imax(imax)
        assert o.cg(which) == imax                 # not converged: stops at imax
        np.testing.assert_allclose(o.get(name)[:n], xs[imax - 1], rtol=0, atol=1e-9 * scale)
    # converged: the loop counter is one more than the updates done, and the
    # stopping test sqrt(u.r)/|b| <= tol fires within one iteration of the
    # textbook's (rounding can move the crossing by one)
    o.set(name, np.ascontiguousarray(x0))
    o.set_cg_imax(200)
    assert o.cg(which) - 1 in (stop, stop + 1)
    np.testing.assert_allclose(o.get(name)[:n], np.linalg.solve(A, b), rtol=0, atol=1e-8 * scale)
    # imax bounds the loop: at imax = stop - 5 it returns imax, unconverged
    o.set(name, np.ascontiguousarray(x0))
    o.set_cg_imax(stop - 5)
    assert o.cg(which) == stop - 5


# ------------------------------------------------------------------ a4 / a7 capacities and Reset
def _icosahedron():
    g = (1 + 5 ** 0.5) / 2
    v = np.array([[0, 1, g], [0, -1, g], [0, 1, -g], [0, -1, -g], [1, g, 0], [-1, g, 0], [1, -g, 0],
                  [-1, -g, 0], [g, 0, 1], [-g, 0, 1], [g, 0, -1], [-g, 0, -1]], dtype=np.float64)
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _dodecahedron():
    g = (1 + 5 ** 0.5) / 2
    v = [[a, b, c] for a in (-1, 1) for b in (-1, 1) for c in (-1, 1)]
    for a in (-1, 1):
        for b in (-1, 1):
            v += [[0, a / g, b * g], [a / g, b * g, 0], [a * g, 0, b / g]]
    v = np.array(v, dtype=np.float64)
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _crowded_carbon():
    """C at the centre, 12 O on an icosahedron at 3.2 A (C-O bonds: the C-O
    BO' reaches 1e-4 at 3.35 A; O-O edges 3.36 A, beyond O-O's 2.54 A), and
    2 H on opposite face normals at 1.0 A (C-H bonded; H-O >= 2.48 A, beyond
    2.39 A; H-H 2.0 A, beyond 1.98 A).  Bond counts by construction: C 14,
    every O and H 1."""
    c = np.array([20.0, 20.0, 20.0])
    ico = _icosahedron()
    face = ico[0] + ico[1] + ico[8]
    face /= np.linalg.norm(face)
    x = np.vstack([c, c + 3.2 * ico, c + 1.0 * face, c - 1.0 * face])
    types = np.array([3] + [2] * 12 + [1, 1], np.int32)      # elements H, O, C
    return x, types


def test_bond_counts_of_the_constructed_geometry(ff):
    x, types = _crowded_carbon()
    el = np.array(["H", "O", "C"])[types - 1]
    cnt = np.zeros(len(x), int)
    for i in range(len(x)):
        for j in range(len(x)):
            d = np.linalg.norm(x[j] - x[i])
            if i != j and d <= 5.0 and sum(ff.bop(el[i], el[j], d)) >= ff.bo_cut:
                cnt[i] += 1
    assert cnt.tolist() == [14] + [1] * 14


def test_bond_capacity_rule_and_reset_offsets(oracle_lib, ffield):
    x, types = _crowded_carbon()
    o = _atoms(oracle_lib, ffield, ["H", "O", "C"], x, types)
    _lists(o)
    # Estimate_Storages: max(2 cnt, 25) per atom
    assert o.get("num_bonds").tolist() == [28] + [25] * 14
    # Reset: bond slots start at the running sum of capacities
    assert o.get("bond_index").tolist() == [0, 28, 53, 78, 103, 128, 153, 178, 203, 228, 253, 278, 303,
                                            328, 353]
    assert (o.get("bond_end") - o.get("bond_index")).tolist() == [14] + [1] * 14
    # Hindex: running index over local H atoms (the last two atoms)
    assert o.get("Hindex").tolist() == [-1] * 13 + [0, 1]
    # each H sees the 12 O within 7.5 A: num_hbonds = (int)max(1.4 * 12, 25) = 25
    # (local atoms that are not H get the minimum too)
    assert o.get("num_hbonds").tolist() == [25] * 15
    assert o.get("hbond_index").tolist() == [0, 25]
    assert (o.get("hbond_end") - o.get("hbond_index")).tolist() == [12, 12]
    # Validate_Lists recomputes the same capacities from the counts it found
    assert o.stage("validate_lists") == 0
    assert o.get("num_bonds").tolist() == [28] + [25] * 14


@pytest.mark.parametrize("nacc,cap", [(18, 25), (19, 26), (20, 28)])
def test_hbond_capacity_rule(oracle_lib, ffield, nacc, cap):
    """One H with nacc O acceptors on a dodecahedron of radius 6 A (inside
    hbond_cut 7.5 A, no O-O or O-H bonds): its hbond count is nacc, and its
    capacity (int)max(1.4 nacc, 25): 18 -> 25 (25.2), 19 -> 26 (26.6),
    20 -> 28."""
    c = np.array([30.0, 30.0, 30.0])
    x = np.vstack([c, c + 6.0 * _dodecahedron()[:nacc]])
    types = np.array([1] + [2] * nacc, np.int32)
    o = _atoms(oracle_lib, ffield, ["H", "O"], x, types)
    _lists(o)
    assert o.get("Hindex").tolist() == [0] + [-1] * nacc
    assert (o.get("hbond_end") - o.get("hbond_index")).tolist() == [nacc]
    assert o.get("num_hbonds")[0] == cap                        # Estimate_Storages
    assert o.get("bond_end").tolist() == o.get("bond_index").tolist()   # no bonds at all
    o.set("num_hbonds", np.array([25] + [0] * nacc, np.int32))
    assert o.stage("validate_lists") == 0
    assert o.get("num_hbonds")[0] == cap                        # Validate_Lists_C


# ------------------------------------------------------------------ a7 sort_bonds
def _h_chain(order):
    """One local H atom (index 0) and five ghost H atoms in a straight chain
    1.2 A apart (H-H bonds reach 1.98 A: only chain neighbours bond).
    order[k] is the atom index of the k-th chain member after the local one."""
    x = np.zeros((6, 3))
    x[0] = [20.0, 20.0, 20.0]
    for k, a in enumerate(order):
        x[a] = [20.0 + 1.2 * (k + 1), 20.0, 20.0]
    return x


@pytest.mark.parametrize("order,marks", [
    # ghosts numbered away from the local atom: one in-order pass reaches the fixed point
    ([1, 2, 3, 4, 5], [0, 1, 2, 3, 4, 5]),
    # numbered towards it: one pass only marks the ghost bonded to the local atom
    ([5, 4, 3, 2, 1], [0, 1000, 1000, 1000, 1000, 1]),
    # mixed numbering
    ([2, 1, 4, 3, 5], [0, 1000, 1, 1000, 1000, 1000]),
])
def test_sort_bonds_single_pass(oracle_lib, ffield, order, marks):
    x = _h_chain(order)
    o = _atoms(oracle_lib, ffield, ["H"], x, np.ones(6, np.int32), n=1, owner=np.zeros(6, np.int32))
    _lists(o)
    # ghosts start at MIN_BONDS capacity; local counts from Estimate_Storages
    assert o.get("num_bonds").tolist() == [25] * 6
    assert o.get("bond_index").tolist() == [0, 25, 50, 75, 100, 125]
    nb = (o.get("bond_end") - o.get("bond_index")).tolist()
    assert nb[0] == 1 and sorted(nb[1:]) == [1, 2, 2, 2, 2]
    assert o.stage("validate_lists") == 0
    assert o.get("bond_mark").tolist() == marks
