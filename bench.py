#!/usr/bin/env python3
"""Benchmark of the nbr + BO + QEq path on one MI355X.

Workload (default): 32,000 rigid waters (96,000 atoms) in a periodic box at
1.0 g/cm^3, seeded lattice placement and orientations, Maxwell velocities at
300 K; NVE at dt = 0.25 fs with the Verlet list rebuilt every 10 steps
(2 A skin); per step the full path runs on the GPU: ghost refresh, QEq
(H build, s/t CG to 1e-6, charges), atom pack, Reset, far list + bonds +
hbonds, list validation and the corrected bond orders.

Prints ONE JSON line: ms/step (lower is better), atom-steps/s, the roofline
of the dominant kernel (HIP events on the library's stream) and the CPU
oracle timed on this machine's host cores on a bounded sample of the same
workload.  The default water96k run adds short legs to the same line: the
mixed-precision QEq on the same box (its solve time against the FP64 solve
and its charge and force errors against FP64 at 1e-10), the batched
clusters512 evaluation and a short serving run.

    python bench.py [--gpus 1] [--steps 200] [--warmup 10] [--config water96k|water1944|water768k]
                    [--mixed] [--cpu-budget-s 20] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from sw_reaxff_amd import _lib, systems  # noqa: E402
from sw_reaxff_amd.lammps import DEFAULT_FFIELD, Lammps, standard_script  # noqa: E402

METRIC = "ms/step (nbr+BO+QEq) at 96k atoms; batched systems/s + p99 at 1e-6 e tol"
HBM_PEAK_GBS = 8000.0
CONFIGS = {
    "water1944": dict(nmol=648, label="648-water box (1,944 atoms)"),
    "water96k": dict(nmol=32000, label="32,000-water box (96,000 atoms)"),
    "water768k": dict(nmol=256000, label="256,000-water box (768,000 atoms)"),
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def device_sync():
    # The path runs on the library's own HIP stream; this is the device-wide
    # barrier (torch is not imported: it carries a second HIP runtime).
    _lib.check(_lib.lib().swr_device_synchronize(), "swr_device_synchronize")


def counts(sysh):
    n, N = sysh.scalar("n"), sysh.scalar("N")
    fi, fe = sysh.get("far_index"), sysh.get("far_end")
    bi, be = sysh.get("bond_index"), sysh.get("bond_end")
    hi, he = sysh.get("hbond_index"), sysh.get("hbond_end")
    return dict(n=n, N=N, verlet=int(sysh.scalar("verlet_total")), far=int((fe - fi).sum()),
                far_local=int((fe[:n] - fi[:n]).sum()),
                bonds=int((be - bi).sum()), bonds_local=int((be[:n] - bi[:n]).sum()), hbonds=int((he - hi).sum()),
                nnz=int(sysh.get("H_numnbrs").sum()))


SLOTS = ("verlet", "bins", "far", "bonds", "reset", "validate", "mark", "bo", "build_H", "panel", "cg_setup", "cg_iter",
         "cg_update", "cg_scalars", "qeq_q", "mixed_inner", "mixed_outer")


def algorithmic_bytes(slot, c, mixed=False):
    """Compulsory HBM bytes of one launch of a kernel slot, from its inputs and outputs."""
    n, N = c["n"], c["N"]
    if slot == "cg_iter":
        # the iteration's SpMV: H (jlist 4 B + val 8 B per entry; 4 B val in
        # the FP32 inner solve), the (s,t) d vector gathered at least once,
        # q written, type and the two row-pointer arrays
        return (8 if mixed else 12) * c["nnz"] + n * (16 + 16 + 4 + 8)
    if slot == "cg_update":
        # the iteration's vector step: 10 (s,t) vectors read (x p u ss w v z r
        # d q), Hdia_inv, 9 written (x p ss v z r u w d)
        return n * (10 * 16 + 8 + 9 * 16)
    if slot == "far":
        # Verlet ids + offsets, packed atoms once, far entries, far_index/end,
        # hbond records, bond-candidate slots (at least one per bond) and counts
        return (4 * c["verlet"] + 8 * (N + 1) + 40 * N + 64 * c["far"] + 8 * N + 16 * c["hbonds"]
                + 4 * c["bonds"] + 8 * N)
    if slot == "bonds":
        # candidate far entries (at least the bonded ones), bond records
        # (56 bond_data + 200 bond_order_data + 8 BO + 16 BOpi + 24 Cdbo*),
        # per-atom sums and marks
        return 64 * c["bonds"] + 304 * c["bonds"] + N * (40 + 8 + 16 + 24 + 4 + 16 + 8)
    if slot == "build_H":
        return 4 * (c["verlet"] * n // max(N, 1)) + 24 * N + 4 * N + 12 * c["nnz"] + n * (8 + 16 + 16 + 80 + 8)
    if slot == "panel":
        # the SpMV panels of H: each entry's column read and its 16-bit list
        # position written, per row its extent and panel-order id, per panel
        # its list (at most one column per entry; counted once per row)
        return 4 * c["nnz"] + 2 * c["nnz"] + n * (4 + 4 + 4) + 4 * n
    if slot == "bo":
        return 8 * N * 20 + c["bonds"] * (56 + 200 + 8 + 16)
    if slot == "verlet":
        return 32 * N + 4 * c["verlet"] + 8 * N
    if slot == "bonded":
        # Atom_Energy: per bond slot of a local atom its partner, BO, BOpi,
        # BO_s and the three coefficients (read and written); per atom the
        # packed record, Delta/Delta_lp/dDelta_lp/Delta_lp_temp, CdDelta (read
        # and written) and its energy rows.  The neighbour terms: per slot the
        # partner, reverse slot, its BOpi and BO.  The chain rule: per slot
        # partner and reverse slot, dBOp, dln_BOp_pi, dln_BOp_pi2, eight C
        # coefficients, the coefficients of both halves, the partner's
        # CdDelta; then per slot the partner's buf_c and dBOp; per atom
        # dDeltap_self, the gradient (read and written twice) and buf_c.
        slots_local = c["bonds_local"]
        return (slots_local * (4 + 8 + 16 + 8 + 48) + n * (40 + 32 + 16 + 16 + 40)
                + c["bonds"] * (8 + 16 + 8) + N * 16
                + c["bonds"] * (8 + 136 + 48 + 8) + c["bonds"] * (8 + 8 + 24) + N * (24 + 48 + 48 + 8))
    if slot == "nonbonded":
        # the local rows' far entries, packed atoms and far_index/end of the
        # local atoms, fCdDelta gradient read and written
        return 64 * c["far_local"] + n * (40 + 8 + 24 + 24)
    return None


def pmc_traffic(kernel_slot, config):
    """Measured HBM bytes per launch of the dominant kernel, from the committed
    rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE, scaled by the copy-kernel
    calibration), or None when no summary covers this kernel and workload."""
    path = os.path.join(ROOT, "profiles", f"pmc_{kernel_slot}_{config}.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(float(line.split(":", 1)[1]))
    except OSError:
        return None
    return None


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def _gpu_forces(sysh):
    """The energy terms and forces of the GPU's last step (the passes the
    harness runs under `forces yes`), for the parity check."""
    L = sysh.L
    for nm in ("swr_atom_energy", "swr_valence_angles", "swr_torsion_angles", "swr_hydrogen_bonds",
               "swr_vdw_coulomb_energy", "swr_add_dbond_forces", "swr_compute_forces"):
        _lib.check(getattr(L, nm)(sysh.h), nm)
    sysh.sync()


def parity_and_cpu_baseline(lmp, sysh, elements, ffield, every, budget_s, samples, one_core=True, oracle_tol=1e-6):
    """The oracle (the FP64 CPU restatement, compiled from oracle/) on the
    GPU's own state, after the timed region: the GPU runs `samples` more
    steps one at a time, and after each the oracle runs the same step from
    the same atoms, Verlet list, capacities and QEq history.  Reports the
    parity of each step (max |dq|, bond orders and coefficients, the lists,
    CG iterations; energies and forces on the first) and the oracle's time
    per step on all host cores and on one core."""
    from oracle import compare, pyoracle  # test-infrastructure checker, only for this leg
    if not os.path.exists(pyoracle.LIB_PATH):
        import subprocess
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, stdout=subprocess.DEVNULL)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = pyoracle.Oracle(ffield, elements, threads=threads, tol=oracle_tol)
    A = compare.atoms_of(sysh)
    n = A["n"]
    o.set_atoms(n, A["x"], A["type"], A["tag"], A["owner"])
    t0 = time.perf_counter()
    o.build_verlet(12.0)
    t_verlet = time.perf_counter() - t0
    t_steps, steps, iters, par = 0.0, 0, [], []
    ef = None
    hist = None
    while steps < samples:
        hist = compare.history_of(sysh)
        lmp.run(1)
        compare.feed_oracle(o, sysh, hist)
        t0 = time.perf_counter()
        its = o.step()
        t_steps += time.perf_counter() - t0
        steps += 1
        git = (int(sysh.scalar("matvecs_s")), int(sysh.scalar("matvecs_t")))
        iters.append((its, git))
        try:
            st = compare.compare_state(sysh, o, 1e-6, far=(steps == 1))    # charges within 1e-6 e
            st["ok"] = True
        except AssertionError as e:
            st = {"ok": False, "error": str(e)[:300]}
        par.append(st)
        if steps == 1:
            _gpu_forces(sysh)
            de, df = compare.compare_energy_forces(sysh, o, A["owner"], n)
            ef = {"max_rel_dE": float(f"{de:.3e}"), "max_rel_dF": float(f"{df:.3e}")}
        if t_steps + t_verlet > budget_s:
            break
    ms = (t_steps / steps + t_verlet / every) * 1e3
    cpu = dict(value=round(ms, 3), unit="ms/step", cores=threads, kind="port",
               sample=f"{steps} oracle steps of the same {n}-atom box on the GPU run's state (+ one Verlet "
                      f"build amortised over {every} steps)",
               nproc=os.cpu_count(), cpu=_cpu_model(), verlet_build_s=round(t_verlet, 3))
    if one_core:
        # the last step again on one core: a Verlet build and one oracle step
        pyoracle.lib().orc_set_threads(1)
        o1 = pyoracle.Oracle(ffield, elements, threads=1)
        compare.feed_oracle(o1, sysh, hist)
        t0 = time.perf_counter()
        o1.build_verlet(12.0)
        t1v = time.perf_counter() - t0
        compare.feed_oracle(o1, sysh, hist)
        t0 = time.perf_counter()
        o1.step()
        t1s = time.perf_counter() - t0
        pyoracle.lib().orc_set_threads(threads)
        cpu["one_core"] = {"value": round((t1s + t1v / every) * 1e3, 3), "unit": "ms/step", "cores": 1,
                           "sample": "1 oracle step of the same box + one Verlet build amortised over "
                                     f"{every} steps", "verlet_build_s": round(t1v, 3)}
    ok = all(p.get("ok") for p in par)
    parity = {"reference": f"FP64 oracle (CPU restatement) at QEq tol {oracle_tol:g}, same atoms, Verlet list, "
                           "capacities and QEq history",
              "steps": steps, "ok": ok,
              "max_abs_dq_e": max((p["dq"] for p in par if p.get("ok")), default=None),
              "max_rel_dBO": max((p["bo_rel"] for p in par if p.get("ok")), default=None),
              "lists_equal": ok, "far_entries_bitwise": par[0].get("far_dq") is not None if par else None,
              "cg_iterations_oracle_gpu": [[list(a), list(b)] for a, b in iters],
              "energies_forces_first_step": ef,
              "failures": [p["error"] for p in par if not p.get("ok")]}
    return cpu, parity


def copy_calibration(mib):
    # A separate mode and process: torch brings its own HIP runtime, so the
    # library is never loaded here.
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "copy_calibration_bytes_written": n * 4,
                      "launches": 5}), flush=True)


def bench_clusters(args, nsteps=None, plain_sample=None):
    """512 non-periodic water/methanol clusters (sizes log-uniform in [50, 3000]):
    one batched evaluation (neighbour list, QEq, lists, BO and the full
    ReaxFF energy + gradient; one launch per stage for all clusters) against
    the plain per-system loop of the same library, both from host arrays to
    results on the device.  Returns the JSON record."""
    from sw_reaxff_amd import batch as B
    count, lo, hi = 512, 50, 3000
    nsteps = max(1, args.steps) if nsteps is None else nsteps
    plain_sample = args.plain_sample if plain_sample is None else plain_sample
    sizes = systems.cluster_sizes(count, lo, hi, seed=args.seed)
    log(f"building {count} clusters, {int(sizes.sum()):,} atoms (seed {args.seed})")
    parts = [systems.cluster(int(k), seed=args.seed + 1 + i) for i, k in enumerate(sizes)]
    natoms = sum(p.n for p in parts)
    bt = B.Batch(tol=1e-6)
    log("batched: warm-up evaluation")
    bt.load(parts)
    bt.evaluate(nonbonded=True, bonded=True)
    bt.sync()
    tune = bt.sys.autotune(reps=3) if args.autotune else None
    lat = []
    for _ in range(nsteps):
        device_sync()
        t0 = time.perf_counter()
        bt.load(parts)
        its = bt.evaluate(nonbonded=True, bonded=True)
        bt.sync()
        lat.append(time.perf_counter() - t0)
    t_batch = float(np.median(lat))
    log(f"batched: {t_batch * 1e3:.1f} ms per batch of {count} (CG iterations max {its})")
    # plain loop: one system at a time through the same library (sampled)
    nplain = min(count, plain_sample)
    one = B.Batch(tol=1e-6)
    one.load([parts[0]])
    one.evaluate(nonbonded=True, bonded=True)
    one.sync()
    device_sync()
    t0 = time.perf_counter()
    for k in range(nplain):
        one.load([parts[k]])
        one.evaluate(nonbonded=True, bonded=True)
        one.sync()
    t_plain = (time.perf_counter() - t0) / nplain
    log(f"plain loop: {t_plain * 1e3:.2f} ms per system ({nplain} systems)")
    sps_batch, sps_plain = count / t_batch, 1.0 / t_plain
    return {
        "metric": "batched systems/s (nbr+BO+QEq + the full ReaxFF energy and gradient, 512 clusters)",
        "value": round(sps_batch, 1),
        "unit": "systems/s", "n_gpus": 1, "steps": len(lat), "warmup": 1,
        "ms_per_step": round(t_batch * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{count} water/methanol clusters, sizes log-uniform [{lo}, {hi}], "
                               f"{natoms:,} atoms, 0.9 g/cm^3, 12 A vacuum padding, QEq tol 1e-6, "
                               "the full ReaxFF energy (vdW, Coulomb, bond, lone pair, over/under-coordination, "
                               "polarization, valence angle, penalty, coalition, torsion, conjugation, hydrogen "
                               "bonds) and its gradient per cluster",
                   "systems": count, "atoms": natoms, "seed": args.seed, "parallelism": "single GPU"},
        "plain_loop_systems_per_s": round(sps_plain, 1), "speedup_vs_plain_loop": round(sps_batch / sps_plain, 2),
        "batch_latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                             "p99": round(float(np.percentile(lat, 99)) * 1e3, 3)},
        "atom_evals_per_s": round(natoms / t_batch, 1), "cg_iters_max": list(its),
        "plain_sample": nplain, "autotune": tune, "host": platform.node(),
    }


# The mixed-precision solve's outer target is tol times this (qeq_mixed.hip).
MIXED_OUTER_FACTOR = 1e-3
QEQ_SLOTS = {False: ("cg_setup", "cg_iter", "cg_update", "cg_scalars", "qeq_q"),
             True: ("mixed_inner", "mixed_outer", "qeq_q")}


def qeq_leg(s, args, mixed, tol, steps):
    """One more run of the same box (its own setup and warmup) with kernel
    timers on: ms/step and the QEq solve's share (H build excluded: the same
    kernel in both forms)."""
    lm = Lammps()
    lm.commands(standard_script(s.elements, ffield=DEFAULT_FFIELD, every=args.every, dt=args.dt, mixed=mixed,
                                tol=tol))
    lm.create_atoms(s)
    lm.setup()
    lm.run(args.warmup)
    h = lm.system
    if args.autotune:
        h.autotune(reps=3)
    h.ktimer_enable(True)
    for slot in SLOTS:
        h.ktimer_read(slot, reset=True)
    its0 = (lm.stat("total_iters_s"), lm.stat("total_iters_t"))
    device_sync()
    t0 = time.perf_counter()
    lm.run(steps)
    device_sync()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    qeq = sum(h.ktimer_read(k)[0] for k in QEQ_SLOTS[mixed]) / steps
    its = ((lm.stat("total_iters_s") - its0[0]) / steps, (lm.stat("total_iters_t") - its0[1]) / steps)
    lm.close()
    return {"form": "mixed" if mixed else "fp64", "tol": tol, "ms_per_step_timed": round(ms, 4),
            "qeq_ms_per_step": round(qeq, 4), "iters_per_step": [round(its[0], 2), round(its[1], 2)]}


def mixed_errors(s, args, steps=None):
    """The mixed-precision run against an FP64 run at a tight tolerance
    (1e-10), stepped side by side from the same atoms: the zero-force NVE
    moves both identically, so every step compares the same positions.  Per
    step: max |dq| (e) and the max error of the non-bonded gradient (the
    forces this path computes, which carry the charges) relative to its
    largest component; the FP64 path at the same tolerance as the mixed one
    runs alongside for comparison."""
    from sw_reaxff_amd import _lib as L
    runs = []
    legs = [(True, args.tol), (False, 1e-10), (False, args.tol)]
    if s.n <= 200_000:   # four systems of the 768,000-atom box do not fit one GPU at once
        legs.append((False, args.tol * MIXED_OUTER_FACTOR))
    for mixed, tol in legs:
        lm = Lammps()
        lm.commands(standard_script(s.elements, ffield=DEFAULT_FFIELD, every=args.every, dt=args.dt,
                                    mixed=mixed, tol=tol))
        lm.create_atoms(s)
        lm.setup()
        runs.append(lm)
    n = s.n
    dq, df, dq64, df64, dqm, dfm = [], [], [], [], [], []
    steps = args.error_steps if steps is None else steps
    for step in range(steps):
        out = []
        for lm in runs:
            lm.run(1)
            h = lm.system
            L.check(h.L.swr_vdw_coulomb_energy(h.h), "swr_vdw_coulomb_energy")
            h.sync()
            out.append((h.get("q")[:n], h.get("fCdDelta").reshape(-1, 4)[:n, :3]))
        (qm, fm), (qd, fd), (q6, f6) = out[:3]
        dq.append(float(np.max(np.abs(qm - qd))))
        df.append(float(np.max(np.abs(fm - fd)) / np.max(np.abs(fd))))
        dq64.append(float(np.max(np.abs(q6 - qd))))
        df64.append(float(np.max(np.abs(f6 - fd)) / np.max(np.abs(fd))))
        if len(out) > 3:
            q9, f9 = out[3]
            dqm.append(float(np.max(np.abs(q9 - qd))))
            dfm.append(float(np.max(np.abs(f9 - fd)) / np.max(np.abs(fd))))
    for lm in runs:
        lm.close()
    matched = None
    if dqm:
        matched = {"tol": args.tol * MIXED_OUTER_FACTOR, "max_abs_dq_e": max(dqm),
                   "median_abs_dq_e": float(np.median(dqm)), "max_rel_dF": max(dfm)}
    return {"steps": steps, "reference": "FP64 QEq at tol 1e-10, same trajectory",
            "max_abs_dq_e": max(dq), "median_abs_dq_e": float(np.median(dq)), "max_rel_dF": max(df),
            "fp64_at_same_tol": {"max_abs_dq_e": max(dq64), "median_abs_dq_e": float(np.median(dq64)),
                                 "max_rel_dF": max(df64)},
            "fp64_at_mixed_outer_target": matched,
            "per_step_max_abs_dq_e": [float(f"{v:.3e}") for v in dq],
            "per_step_max_rel_dF": [float(f"{v:.3e}") for v in df]}


def mixed_compare(s, args, qeq_mixed, ksteps, error_steps):
    """The mixed-precision QEq against the FP64 one on box s: its charge and
    force errors against FP64 at 1e-10 over error_steps side-by-side steps,
    and its solve time (qeq_mixed ms/step) against the FP64 solve at the
    same tolerance and at the mixed solve's own outer target (the FP64
    tolerance that reaches the same accuracy), each timed in a run of its
    own."""
    errors = mixed_errors(s, args, error_steps) if error_steps > 0 else None
    if errors:
        log(f"mixed vs FP64: {errors}")
    legs = [qeq_leg(s, args, False, args.tol, ksteps),
            qeq_leg(s, args, False, args.tol * MIXED_OUTER_FACTOR, ksteps)]
    cmp = {"mixed_qeq_ms_per_step": round(qeq_mixed, 4), "fp64": legs,
           "speedup_vs_fp64_same_tol": round(legs[0]["qeq_ms_per_step"] / qeq_mixed, 3),
           "speedup_vs_fp64_at_outer_target": round(legs[1]["qeq_ms_per_step"] / qeq_mixed, 3)}
    log(f"QEq mixed vs FP64: {cmp}")
    return errors, cmp


def bench_serve(args, rates=None, requests=None, reps=None, pool=None):
    """A Poisson stream of evaluation requests (sizes log-uniform in [20, 20000],
    half periodic water boxes, half clusters, 30% follow-on MD steps of
    resident systems) through the continuously batching server at a sweep of
    arrival rates; the value is the highest throughput whose p99 latency stays
    within 50 ms.  The plain loop is the same server with one worker and one
    request per evaluation, all requests queued at the start.  Returns the
    JSON record."""
    from sw_reaxff_amd import serve
    from sw_reaxff_amd.batch import Batch
    mixed = not args.fp64
    lanes = [Batch(tol=args.tol, mixed=mixed) for _ in range(args.lanes)]
    serve.warm_lanes(lanes, args.max_batch_atoms)
    tune = None
    if args.autotune:
        # the shapes tuned on one lane's warm-up batch, given to every lane
        tune = lanes[0].sys.autotune(reps=3)
        for bt in lanes[1:]:
            for fam, rec in tune.items():
                if isinstance(rec, dict) and "chosen" in rec:
                    bt.sys.set_shape(fam, rec["chosen"])
    warm = serve.make_requests(24, rate=1e4, seed=args.seed + 1, hi=3000)
    serve.Server(lanes=args.lanes, max_batch_atoms=args.max_batch_atoms, batches=lanes, tol=args.tol,
                 mixed=mixed, resident=args.resident).run(warm)
    rates = args.rates if rates is None else rates
    requests = args.requests if requests is None else requests
    nreps = args.reps if reps is None else reps
    pool = args.pool if pool is None else pool
    runs = []
    # the stream's systems are generated once (drawn from `pool` distinct
    # systems) and replayed at every rate
    log(f"serving: {requests} requests from {pool} distinct systems")
    base = serve.make_requests(requests, rate=1.0, seed=args.seed, pool=pool)
    for rate in rates:
        reps = []
        for rep in range(nreps):
            reqs = serve.at_rate(base, rate)
            t0 = time.perf_counter()
            srv = serve.Server(lanes=args.lanes, max_batch_atoms=args.max_batch_atoms, batches=lanes, tol=args.tol,
                               mixed=mixed, resident=args.resident)
            srv.run(reqs)
            wall = time.perf_counter() - t0
            summ = serve.summary(reqs, wall)
            summ["rate"] = rate
            summ["lane_busy_s"] = {k: round(v, 3) for k, v in srv.busy.items()}
            summ["wall_s"] = round(wall, 3)
            reps.append(summ)
            log(f"rate {rate:.0f}/s (run {rep + 1}/{nreps}): {summ}")
        # a rate counts with its slowest repetition: lowest throughput, highest p99
        worst = min(reps, key=lambda r: r["systems_per_s"])
        worst = dict(worst, latency_ms=max((r["latency_ms"] for r in reps), key=lambda l: l["p99"]),
                     completed=min(r["completed"] for r in reps), repetitions=len(reps))
        runs.append(worst)
    # plain loop: one worker, batches of one, every request queued at t = 0
    reqs = serve.at_rate(base[:args.plain_sample], 1e9)
    t0 = time.perf_counter()
    serve.Server(lanes=1, max_batch_atoms=1, batches=[lanes[0]], tol=args.tol, mixed=mixed,
                 resident=args.resident).run(reqs)
    wall = time.perf_counter() - t0
    plain = len(reqs) / wall
    log(f"plain loop: {plain:.1f} systems/s ({len(reqs)} requests)")
    ok = [r for r in runs if r["latency_ms"]["p99"] <= 50.0 and r["completed"] == r["requests"]]
    best = max(ok, key=lambda r: r["systems_per_s"]) if ok else max(runs, key=lambda r: r["systems_per_s"])
    out = {
        "metric": "serving systems/s at p99 <= 50 ms (nbr+BO+QEq, Poisson arrivals)", "value": best["systems_per_s"],
        "unit": "systems/s", "n_gpus": 1, "steps": requests, "warmup": len(warm),
        "ms_per_step": best["latency_ms"]["p50"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64+f32" if mixed else "f64", "data": "synthetic",
        "config": {"workload": f"{requests} requests per rate ({pool} distinct systems), Poisson arrivals, "
                               "sizes log-uniform [20, 20000], "
                               "50% periodic water / 50% water-methanol clusters, 30% follow-on MD steps with "
                               f"resident s/t, 200 GB HBM admission, "
                               f"{'mixed-precision ' if mixed else 'FP64 '}QEq tol {args.tol:g}",
                   "lanes": args.lanes, "max_batch_atoms": args.max_batch_atoms, "seed": args.seed,
                   "residents": args.resident,
                   "parallelism": "single GPU"},
        "best_rate": best["rate"], "p99_within_50ms": bool(ok), "latency_ms": best["latency_ms"],
        "runs": runs, "plain_loop_systems_per_s": round(plain, 1),
        "speedup_vs_plain_loop": round(best["systems_per_s"] / plain, 2), "autotune": tune,
        "host": platform.node(),
    }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="water96k", choices=sorted(CONFIGS) + ["clusters512", "serve"])
    ap.add_argument("--requests", type=int, default=4000, help="serve: requests in the stream (per rate)")
    ap.add_argument("--pool", type=int, default=400, help="serve: distinct systems the requests draw from")
    ap.add_argument("--rates", type=float, nargs="+", default=[1000, 1500, 2000, 2400, 2600, 2800],
                    help="serve: Poisson arrival rates (1/s) to sweep")
    ap.add_argument("--lanes", type=int, default=4, help="serve: concurrent batch workers")
    ap.add_argument("--reps", type=int, default=2, help="serve: runs of the stream per rate (the slowest counts)")
    ap.add_argument("--max-batch-atoms", type=int, default=200_000, help="serve: atoms per batch at most")
    ap.add_argument("--plain-sample", type=int, default=128,
                    help="clusters timed one at a time for the plain-loop comparison")
    ap.add_argument("--mixed", action="store_true", help="mixed-precision QEq")
    ap.add_argument("--fp64", action="store_true", help="serve: FP64 QEq instead of the mixed-precision form")
    ap.add_argument("--resident", default="batched", choices=["batched", "device"],
                    help="serve: follow-on steps in the batches (warm start in device memory) or on each "
                         "resident's own device system")
    ap.add_argument("--tol", type=float, default=1e-6, help="QEq tolerance (the reference's: 1e-6)")
    ap.add_argument("--error-steps", type=int, default=200,
                    help="mixed: steps of the side-by-side comparison with an FP64 run (0: skip)")
    ap.add_argument("--leg-error-steps", type=int, default=20,
                    help="water96k default run: side-by-side steps of its mixed-precision leg's error check")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="launch with the default shapes instead of the autotuner's")
    ap.add_argument("--no-legs", dest="legs", action="store_false",
                    help="water96k: skip the short clusters512 and serving runs whose results join the line")
    ap.add_argument("--serve-rate", type=float, default=1500.0, help="arrival rate of the short serving run")
    ap.add_argument("--serve-requests", type=int, default=2000, help="requests of the short serving run")
    ap.add_argument("--serve-pool", type=int, default=150, help="distinct systems of the short serving run")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--every", type=int, default=10)
    ap.add_argument("--dt", type=float, default=0.25)
    ap.add_argument("--cpu-budget-s", type=float, default=20.0)
    ap.add_argument("--parity-steps", type=int, default=3,
                    help="GPU steps after the timed region checked against the oracle (also the CPU sample)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain copy of MIB MiB (read MIB, write MIB) five times: the known-size "
                         "kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this path has no multi-GPU form (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    if args.config == "clusters512":
        print(json.dumps(bench_clusters(args)), flush=True)
        return
    if args.config == "serve":
        print(json.dumps(bench_serve(args)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this path has no multi-GPU form (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    cfg = CONFIGS[args.config]
    log(f"building {cfg['label']} seed {args.seed}")
    s = systems.water_box(cfg["nmol"], seed=args.seed)
    lmp = Lammps()
    lmp.commands(standard_script(s.elements, ffield=DEFAULT_FFIELD, every=args.every, dt=args.dt,
                                 mixed=args.mixed, tol=args.tol))
    lmp.create_atoms(s)
    log("setup (first neighbour build, capacity estimate, QEq from zero history)")
    lmp.setup()
    sysh = lmp.system
    log(f"warmup {args.warmup} steps")
    lmp.run(args.warmup)
    # launch shapes chosen by the library's autotuner on this box, before the
    # timed region (the choices and per-shape times go into the line)
    tune = None
    if args.autotune:
        tune = sysh.autotune(reps=5)
        log(f"autotune: {tune}")
        lmp.run(2)
    its0 = (lmp.stat("total_iters_s"), lmp.stat("total_iters_t"))
    device_sync()
    # the headline: whole steps, no per-kernel event records in the stream
    log(f"timed {args.steps} steps")
    t0 = time.perf_counter()
    lmp.run(args.steps)
    device_sync()
    elapsed = time.perf_counter() - t0
    ms_step = elapsed * 1e3 / args.steps
    its = ((lmp.stat("total_iters_s") - its0[0]) / args.steps, (lmp.stat("total_iters_t") - its0[1]) / args.steps)
    # kernel table and roofline: the same run continued with HIP-event timers
    # around every kernel (they add a few us per launch, so they stay out of
    # the headline)
    ksteps = max(10, min(args.steps, 50))
    sysh.ktimer_enable(True)
    for slot in SLOTS:
        sysh.ktimer_read(slot, reset=True)
    work0 = lmp.stat("cg_work_launches")
    log(f"kernel timers over {ksteps} further steps")
    lmp.run(ksteps)
    device_sync()
    cg_work = int(lmp.stat("cg_work_launches") - work0)
    kt = {}
    for slot in SLOTS:
        ms, cnt = sysh.ktimer_read(slot)
        if cnt:
            kt[slot] = (ms, cnt)
    sysh.ktimer_enable(False)
    c = counts(sysh)
    n = c["n"]
    log(f"ms/step {ms_step:.3f}; CG iterations s {its[0]:.1f} t {its[1]:.1f}; kernels {kt}")
    # dominant kernel and its roofline.  Iteration launches enqueued after
    # both solves converged exit at once; only the launches that did an
    # iteration's work count for the per-launch average.
    def work_launches(slot, cnt):
        return max(1, cg_work) if slot in ("cg_iter", "cg_update") else cnt

    dom = max(kt, key=lambda k: kt[k][0])
    dom_ms, dom_cnt = kt[dom]
    launches = work_launches(dom, dom_cnt)
    per_launch_bytes = algorithmic_bytes(dom, c, args.mixed)
    avg_ms = dom_ms / launches
    achieved = per_launch_bytes / (avg_ms * 1e-3) / 1e9 if per_launch_bytes else None
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1) if achieved else None,
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
                "traffic": pmc_traffic(dom, args.config),
                "bytes_per_launch": per_launch_bytes, "avg_launch_us": round(avg_ms * 1e3, 2),
                "launches": launches}
    kernels = {}
    for k, (ms, cnt) in kt.items():
        b = algorithmic_bytes(k, c, args.mixed)
        wl = work_launches(k, cnt)
        kernels[k] = {"ms_per_step": round(ms / ksteps, 4), "launches": cnt,
                      "avg_us": round(ms * 1e3 / wl, 2),
                      "gbs": round(b * wl / (ms * 1e-3) / 1e9, 1) if b else None}
    # Widening beyond the metric's path: the non-bonded pass
    # (vdW_Coulomb_Energy_Full_C) over the last step's far list, timed alone
    sysh.ktimer_enable(True)
    sysh.ktimer_read("nonbonded", reset=True)
    for _ in range(10):
        from sw_reaxff_amd import _lib as _L
        _L.check(sysh.L.swr_vdw_coulomb_energy(sysh.h), "swr_vdw_coulomb_energy")
    device_sync()
    nb_ms, nb_cnt = sysh.ktimer_read("nonbonded")
    sysh.ktimer_enable(False)
    nb_avg = nb_ms / max(1, nb_cnt)
    nb_b = algorithmic_bytes("nonbonded", c)
    nonbonded = {"kernel": "k_vdw_coulomb", "avg_us": round(nb_avg * 1e3, 2), "launches": nb_cnt,
                 "bytes_per_launch": nb_b, "achieved_gbs": round(nb_b / (nb_avg * 1e-3) / 1e9, 1),
                 "frac": round(nb_b / (nb_avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                 "energies_kcal_mol": [float(v) for v in sysh.get("nonbonded_energy")[:2]]}
    log(f"non-bonded pass: {nonbonded}")
    # and the bond/atom energies with the bond-order chain rule of the
    # gradient (Atom_Energy + Add_All_dBond_to_Forces) over the same bonds
    sysh.ktimer_enable(True)
    for slot in ("atom_energy", "valence", "torsion", "hbonds", "dbond"):
        sysh.ktimer_read(slot, reset=True)
    for _ in range(10):
        _L.check(sysh.L.swr_atom_energy(sysh.h), "swr_atom_energy")
        _L.check(sysh.L.swr_valence_angles(sysh.h), "swr_valence_angles")
        _L.check(sysh.L.swr_torsion_angles(sysh.h), "swr_torsion_angles")
        _L.check(sysh.L.swr_hydrogen_bonds(sysh.h), "swr_hydrogen_bonds")
        _L.check(sysh.L.swr_add_dbond_forces(sysh.h), "swr_add_dbond_forces")
    device_sync()
    ae_ms, ae_cnt = sysh.ktimer_read("atom_energy")
    hb_ms, hb_cnt = sysh.ktimer_read("hbonds")
    va_ms, va_cnt = sysh.ktimer_read("valence")
    to_ms, to_cnt = sysh.ktimer_read("torsion")
    db_ms, db_cnt = sysh.ktimer_read("dbond")
    sysh.ktimer_enable(False)
    be_b = algorithmic_bytes("bonded", c)
    be_avg = (ae_ms + db_ms) / max(1, ae_cnt)
    bonded = {"kernels": "k_atom_energy, k_atom_energy_nbr, k_sum_rows, k_dbond_own, k_dbond_nbr",
              "avg_us": round(be_avg * 1e3, 2), "atom_energy_us": round(ae_ms * 1e3 / max(1, ae_cnt), 2),
              "dbond_us": round(db_ms * 1e3 / max(1, db_cnt), 2), "launches": ae_cnt, "bytes_per_pass": be_b,
              "achieved_gbs": round(be_b / (be_avg * 1e-3) / 1e9, 1),
              # event spans from the first to the last launch of each pass,
              # which read longer than the kernels' own durations (rocprof)
              "valence": {"kernels": "k_valence, k_va_fold, k_sum_rows",
                          "event_span_us": round(va_ms * 1e3 / max(1, va_cnt), 2),
                          "energies_kcal_mol": [float(v) for v in sysh.get("valence_energy")]},
              "torsion": {"kernels": "k_torsion, k_va_fold, k_to_slot_fold, k_sum_rows",
                          "event_span_us": round(to_ms * 1e3 / max(1, to_cnt), 2),
                          "energies_kcal_mol": [float(v) for v in sysh.get("torsion_energy")]},
              "hbonds": {"kernels": "k_hb, k_hb_fold, k_sum_rows",
                         "event_span_us": round(hb_ms * 1e3 / max(1, hb_cnt), 2),
                         "e_hb_kcal_mol": float(sysh.get("hbond_energy")[0])},
              "energies_kcal_mol": dict(zip(("e_bond", "e_lp", "e_ov", "e_un", "e_pol"),
                                            [float(v) for v in sysh.get("bonded_energy")]))}
    log(f"bonded pass: {bonded}")
    cpu, parity = None, None
    if not args.no_cpu:
        log(f"oracle on the GPU's state: parity and CPU baseline (budget {args.cpu_budget_s:.0f} s)")
        # the mixed-precision charges are checked against the converged FP64 solution
        cpu, parity = parity_and_cpu_baseline(lmp, sysh, s.elements, DEFAULT_FFIELD, args.every,
                                              args.cpu_budget_s, args.parity_steps, one_core=not args.mixed,
                                              oracle_tol=1e-10 if args.mixed else args.tol)
        log(f"CPU {cpu['value']} ms/step on {cpu['cores']} threads; {cpu.get('one_core')}; parity {parity}")
    errors, qeq_cmp = None, None
    if args.mixed:
        qeq_mixed = sum(kt[k][0] for k in QEQ_SLOTS[True] if k in kt) / ksteps
        lmp.close()   # the comparisons hold several systems of this size at once
        errors, qeq_cmp = mixed_compare(s, args, qeq_mixed, ksteps, args.error_steps)
    # the metric's second half (batched systems/s, serving p50/p99) as extra
    # keys of the same line: short runs of the clusters512 and serve workloads;
    # and the mixed-precision QEq on the same box (its own run, then the FP64
    # comparison runs and a short side-by-side error check)
    legs = {}
    if args.legs and not args.mixed:
        lmp.close()
        log(f"mixed-precision QEq on the same box ({args.leg_error_steps} side-by-side error steps)")
        mq = qeq_leg(s, args, True, args.tol, ksteps)
        errors, qeq_cmp = mixed_compare(s, args, mq["qeq_ms_per_step"], ksteps, args.leg_error_steps)
        qeq_cmp["mixed"] = mq
        cl = bench_clusters(args, nsteps=3, plain_sample=64)
        legs["clusters512"] = {k: cl[k] for k in ("value", "unit", "ms_per_step", "batch_latency_ms",
                                                  "plain_loop_systems_per_s", "speedup_vs_plain_loop",
                                                  "atom_evals_per_s", "cg_iters_max", "plain_sample")}
        legs["clusters512"]["workload"] = cl["config"]["workload"]
        sv = bench_serve(args, rates=[args.serve_rate], requests=args.serve_requests, reps=1,
                         pool=args.serve_pool)
        legs["serving"] = {k: sv[k] for k in ("value", "unit", "best_rate", "latency_ms", "p99_within_50ms",
                                              "plain_loop_systems_per_s", "speedup_vs_plain_loop")}
        legs["serving"]["workload"] = sv["config"]["workload"]
        legs["serving"]["lanes"] = sv["config"]["lanes"]
        legs["serving"]["residents"] = sv["config"]["residents"]
    out = {
        "metric": METRIC, "value": round(ms_step, 4), "unit": "ms/step", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if not args.mixed else "f64+f32", "data": "synthetic",
        "config": {"workload": cfg["label"] + f", NVE 300 K dt {args.dt} fs, rebuild every {args.every} steps, "
                               f"skin 2 A, QEq tol {args.tol:g}" + (", mixed-precision QEq" if args.mixed else ""),
                   "atoms": n, "atoms_with_ghosts": c["N"], "seed": args.seed, "parallelism": "single GPU"},
        "atom_steps_per_s": round(n / (ms_step * 1e-3), 1),
        "cg_iters_per_step": [round(its[0], 2), round(its[1], 2)],
        "counts": c, "kernels": kernels, "roofline": roofline, "cpu_baseline": cpu, "parity_vs_oracle": parity,
        "autotune": tune,
        "nonbonded_pass": nonbonded, "bonded_pass": bonded, "mixed_errors_vs_fp64": errors, "mixed_qeq_vs_fp64": qeq_cmp,
        "speedup_vs_cpu": round(cpu["value"] / ms_step, 2) if cpu else None,
        "speedup_vs_cpu_one_core": round(cpu["one_core"]["value"] / ms_step, 2) if cpu and cpu.get("one_core") else None,
        **legs,
        "host": platform.node(),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
