"""SpMV panels of the QEq matrix.

The local rows of H are ordered along a Morton curve over the neighbour
list's bins, cut into panels of at most 64 rows of one system, and each
panel lists its distinct columns; every H entry holds its column's position
in that list, and the SpMV gathers x from the list in LDS.  A panel with
more columns than the list holds (3072) gathers from global memory.

Checked here: the lists index exactly their panels' columns; the order is
a permutation keeping each system's rows at the system's own row range; and
the charges and iteration counts are the same bits with the lists on and
off (the two paths sum each row's entries in the same order), for the FP64
solve, the mixed-precision one, a batch of clusters, and a compressed box
whose panels overflow their lists.
"""
import numpy as np
import pytest

from sw_reaxff_amd import _lib, systems

from .parity import harness, water_c0

pytestmark = pytest.mark.gpu


def _check_lists(h, row_ranges):
    n = h.scalar("n")
    prow, pbeg, pcnt = h.get("panel_rows"), h.get("panel_begin"), h.get("panel_count")
    pcol = h.get("panel_columns").reshape(-1, _lib.PANEL_MAXC)
    f, nn, jl, jloc = h.get("H_firstnbr"), h.get("H_numnbrs"), h.get("H_jlist"), h.get("H_jloc")
    assert np.array_equal(np.sort(prow), np.arange(n))
    for a, b in row_ranges:   # each system's rows stay at its own range
        assert np.array_equal(np.sort(prow[a:b]), np.arange(a, b))
    starts = {a for a, _ in row_ranges}
    assert pbeg[0] == 0 and pbeg[-1] == n and np.all(np.diff(pbeg) >= 1) and np.all(np.diff(pbeg) <= 64)
    ends = {b for _, b in row_ranges}
    for p in range(len(pbeg) - 1):   # no panel crosses a system boundary
        assert not any(pbeg[p] < e < pbeg[p + 1] for e in ends), p
    assert starts <= set(pbeg.tolist())
    nlist, entries = 0, 0
    for p in np.flatnonzero(pcnt >= 0):
        m = int(pcnt[p])
        cols = pcol[p, :m]
        assert len(np.unique(cols)) == m, p
        used = []
        for i in prow[pbeg[p]:pbeg[p + 1]]:
            sl = slice(f[i], f[i] + ((nn[i] + 3) & ~3))   # padding entries included
            assert np.all(jloc[sl] < m)
            assert np.array_equal(cols[jloc[sl]], jl[sl]), (p, i)
            used.append(jl[sl])
        assert np.array_equal(np.unique(np.concatenate(used)), np.sort(cols)), p
        nlist += m
        entries += sum(len(u) for u in used)
    return nlist, entries, int((pcnt < 0).sum())


def _resolve(h, hist):
    h.put("s_hist", hist[0])
    h.put("t_hist", hist[1])
    h.call("qeq_init_matvec")
    its = h.qeq_cg()
    h.call("calculate_q")
    h.sync()
    return h.get("q")[: h.scalar("n")].copy(), its


def test_panel_lists_index_their_columns(gpu_lib, ffield):
    lmp = harness(water_c0(seed=51), ffield, tol=1e-10)
    h = lmp.system
    nlist, entries, overflow = _check_lists(h, [(0, h.scalar("n"))])
    print(f"[panels] {h.get('panel_count').size} panels, {nlist} listed columns for {entries} entries, "
          f"{overflow} without a list", flush=True)
    assert overflow == 0
    assert nlist < 0.25 * entries   # compact panels: most entries share columns


@pytest.mark.parametrize("mixed", [False, True])
def test_panels_on_off_same_bits(gpu_lib, ffield, mixed):
    lmp = harness(water_c0(seed=52), ffield, tol=1e-10 if not mixed else 1e-6, mixed=mixed)
    h = lmp.system
    hist = (h.get("s_hist").copy(), h.get("t_hist").copy())
    q_on, its_on = _resolve(h, hist)
    h.set_shape("panels", "off")   # from the next H build on
    assert h.shape("panels") == "off"
    q_off, its_off = _resolve(h, hist)
    assert np.all(h.get("panel_count") < 0)
    h.set_shape("panels", "on")
    print(f"[panels] mixed={mixed}: iterations on {its_on} off {its_off}", flush=True)
    assert its_on == its_off
    assert np.array_equal(q_on, q_off)


def test_batch_panels_keep_systems_apart(gpu_lib):
    from sw_reaxff_amd import batch as B
    parts = [systems.cluster(int(k), seed=60 + i) for i, k in enumerate([120, 800, 45, 300, 1500, 64, 65])]
    bt = B.Batch(tol=1e-10)
    P = bt.load(parts)
    its = bt.evaluate()
    bt.sync()
    h = bt.sys
    rng = [(int(P.row_off[k]), int(P.row_off[k + 1])) for k in range(len(parts))]
    _check_lists(h, rng)
    q_on = np.concatenate(bt.charges()).copy()
    h.set_shape("panels", "off")
    bt.load(parts)
    its_off = bt.evaluate()
    bt.sync()
    q_off = np.concatenate(bt.charges())
    print(f"[panels] batch of {len(parts)}: iterations on {its} off {its_off}", flush=True)
    assert its == its_off and np.array_equal(q_on, q_off)


def test_dense_box_panels_overflow_to_global(gpu_lib, ffield):
    """A water box compressed to ~2x density: its largest panels touch more
    columns than their lists hold, and those rows gather from global memory."""
    s = systems.water_box(2000, seed=53)
    f = 0.8
    c = (s.lo + s.hi) / 2
    s = systems.System(x=(s.x - c) * f + c, v=s.v, type=s.type, tag=s.tag, elements=s.elements,
                       lo=(s.lo - c) * f + c, hi=(s.hi - c) * f + c, periodic=s.periodic)
    lmp = harness(s, ffield, tol=1e-10)
    h = lmp.system
    nlist, entries, overflow = _check_lists(h, [(0, h.scalar("n"))])
    print(f"[panels] dense: {overflow} of {h.get('panel_count').size} panels without a list", flush=True)
    assert overflow > 0
    hist = (h.get("s_hist").copy(), h.get("t_hist").copy())
    q_on, its_on = _resolve(h, hist)
    h.set_shape("panels", "off")
    q_off, its_off = _resolve(h, hist)
    assert np.all(h.get("panel_count") < 0)
    assert its_on == its_off and np.array_equal(q_on, q_off)
