"""Serving loop: evaluation requests of independent systems, continuously batched.

Requests arrive over time; each is either a new system to evaluate once
(neighbour list, QEq charges, far/bond/hbond lists, bond orders, the
non-bonded and bond/atom energies and their gradient) or a follow-on MD
step of a resident system.  The
server:

  * runs `lanes` workers, each owning one device system (its own HIP stream).
    A free worker takes the queued requests in arrival order that fit the
    HBM budget and the batch size cap and evaluates them as ONE batch
    (Batch: one launch per stage for all of them); requests that arrive
    while a batch is on the GPU wait for the next free worker, and the
    workers' batches overlap on the device (ctypes releases the GIL during
    library calls);
  * resident systems (resident="batched", the default): a follow-on step
    is one velocity-Verlet step of fix nve with the ReaxFF forces: the
    half kick and drift with the forces its last evaluation computed
    (nve_initial), the evaluation of the new positions in the next batch
    like any other request (its QEq solve warm-started from s/t histories
    kept in device memory between its steps: HistStore, saved from the
    batch that evaluated its last step, loaded into the batch of its next
    one; all five slots hold the solution after its first evaluation), and
    the final half kick with the new forces (nve_final).  The few hundred
    bytes per atom of positions, velocities and forces go with the batch's
    host packing; the Verlet list is built per batch, whose composition
    changes with every batch.  Steps of one system run in order, one per
    batch;
  * resident="device": each resident system instead keeps its own device
    system (positions, velocities, forces, Verlet list with skin, s/t
    history) set up at its first follow-on step from the histories its
    batched solve left, and each further step runs there with the harness'
    fix nve and forces (one system per launch);
  * admits work by estimated HBM footprint against a budget (bytes per atom
    dominated by the Verlet and far lists).

Latency is measured from a request's arrival to its charges being on the
host.  Nothing here changes a result: a system's charges are the ones it
gets evaluated alone (tests/test_gpu_serve.py).
"""
from __future__ import annotations

import math
import threading
import time
from dataclasses import dataclass, replace

import numpy as np

from . import systems
from .batch import Batch, CUTGHOST, ELEMENTS
from .lammps import DEFAULT_FFIELD, Lammps, standard_script

# Per-atom device bytes (local + ghost): ~700 Verlet ids (4 B) + ~420 far
# entries (64 B) + bonds, H, CG vectors.
BYTES_PER_ATOM = 32 * 1024


@dataclass
class Request:
    rid: int
    arrival: float                  # seconds from the start of the run
    kind: str                       # "new" | "step"
    system: systems.System | None = None
    resident: int = -1              # resident system id (kind "step", or a new system kept resident)
    # filled in by the server
    done: float = 0.0
    q: np.ndarray | None = None
    iters: tuple = ()
    lane: int = -1
    batch_size: int = 0
    step_index: int = 0             # kind "step": which step of its resident system this was
    hist: tuple | None = None       # resident's first evaluation: the s/t histories it left
    energy: float | None = None     # the full ReaxFF energy of the evaluation (kcal/mol)
    f: np.ndarray | None = None     # residents' evaluations: the forces on the local atoms (kcal/mol/A)

    @property
    def latency(self):
        return self.done - self.arrival


def footprint(s: systems.System) -> int:
    """Estimated device bytes of a system (ghost shell included when periodic)."""
    n = s.n
    if s.periodic:
        L = float(np.min(s.hi - s.lo))
        n = int(n * ((L + 2 * CUTGHOST) / L) ** 3)
    return n * BYTES_PER_ATOM


def make_system(kind: str, natoms: int, seed: int) -> systems.System:
    if kind == "water":
        # the periodic box must be at least one ghost cutoff wide (L >= 12 A: >= 58 waters)
        return systems.water_box(max(64, natoms // 3), seed=seed)
    return systems.cluster(max(3, natoms), seed=seed)


def make_requests(count: int, rate: float, seed: int, lo: int = 20, hi: int = 20000, frac_step: float = 0.3,
                  frac_resident: float = 0.2, pool: int = 0):
    """Poisson arrivals at `rate` per second; sizes log-uniform in [lo, hi];
    half periodic water boxes, half water/methanol clusters; a fraction of
    requests are follow-on MD steps of a resident system (new systems become
    resident with probability frac_resident).  With pool > 0 the new
    requests draw their systems from `pool` distinct systems generated first
    (same size and kind distribution; building thousands of distinct
    clusters in Python takes minutes), otherwise each is generated anew."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=count))
    systems_pool = []
    for _ in range(pool):
        size = int(math.exp(rng.uniform(math.log(lo), math.log(hi))))
        kind = "water" if rng.random() < 0.5 else "cluster"
        systems_pool.append(make_system(kind, size, seed=int(rng.integers(1 << 30))))
    reqs, residents = [], []
    for k in range(count):
        if residents and rng.random() < frac_step:
            reqs.append(Request(rid=k, arrival=float(t[k]), kind="step", resident=int(rng.choice(residents))))
            continue
        if systems_pool:
            s = systems_pool[int(rng.integers(len(systems_pool)))]
        else:
            size = int(math.exp(rng.uniform(math.log(lo), math.log(hi))))
            kind = "water" if rng.random() < 0.5 else "cluster"
            s = make_system(kind, size, seed=int(rng.integers(1 << 30)))
        r = Request(rid=k, arrival=float(t[k]), kind="new", system=s)
        if rng.random() < frac_resident:
            r.resident = len(residents)
            residents.append(r.resident)
        reqs.append(r)
    return reqs


def at_rate(reqs: list[Request], rate: float, unit_rate: float = 1.0) -> list[Request]:
    """Fresh copies of requests generated at `unit_rate`, arriving at `rate`
    (same systems and kinds, arrival times scaled)."""
    f = unit_rate / rate
    return [Request(rid=r.rid, arrival=r.arrival * f, kind=r.kind, system=r.system, resident=r.resident)
            for r in reqs]


# force->ftm2v in LAMMPS real units: (kcal/mol/A) / amu -> A/fs^2
FTM2V = 1.0 / 48.88821291 / 48.88821291


def _wrap(x, s: systems.System):
    """Periodic positions wrapped into [lo, hi) as Domain::pbc does."""
    if not s.periodic:
        return x
    L = s.hi - s.lo
    x = np.where(x < s.lo, x + L, x)
    over = x >= s.hi
    return np.where(over, np.maximum(x - L, s.lo), x)


def nve_initial(s: systems.System, f: np.ndarray, dt: float) -> systems.System:
    """FixNVESunway::initial_integrate (fix_nve_sw64.c:25-99, per-type mass
    branch): v += dtf/m f, x += dt v with dtf = dt/2 ftm2v
    (fix_nve_sunway.cpp:56-57), then periodic wrapping."""
    dtf = 0.5 * dt * FTM2V
    v = s.v + (dtf / s.masses)[:, None] * f
    return replace(s, x=_wrap(s.x + dt * v, s), v=v)


def nve_final(s: systems.System, f: np.ndarray, dt: float) -> systems.System:
    """FixNVESunway::final_integrate (fix_nve_sw64.c:102-173): v += dtf/m f."""
    return replace(s, v=s.v + (0.5 * dt * FTM2V / s.masses)[:, None] * f)


class _ResidentState:
    """A resident system between its steps (resident="batched"): positions,
    velocities (after the last step's final kick) and the forces at those
    positions on the host, its QEq warm start in device memory."""

    def __init__(self, s: systems.System, f: np.ndarray, store):
        self.s, self.f, self.store = s, f, store
        self.steps = 0
        self.busy = False                # a step of it is in a batch on the GPU


class _Resident:
    """A system kept on its own device system between its MD steps
    (resident="device").  The device system is created and set up on first
    use (first() or the first step()), warm-started from `hist` (the s/t
    histories of its first evaluation) when given."""

    def __init__(self, s: systems.System, ffield, tol, mixed, hist=None):
        self.lock = threading.Lock()
        self.s, self.ffield, self.tol, self.mixed, self.hist = s, ffield, tol, mixed, hist
        self.lmp = None
        self.n = s.n
        self.bytes = footprint(s)
        self.steps = 0

    def _ensure(self):
        if self.lmp is None:
            lmp = Lammps()
            lmp.commands(standard_script(self.s.elements, ffield=self.ffield, tol=self.tol, mixed=self.mixed,
                                         forces=True))
            lmp.create_atoms(self.s)
            if self.hist is not None:
                lmp.set_history(*self.hist)
            lmp.setup()
            self.lmp = lmp

    def first(self):
        self._ensure()
        return self.q()

    def step(self):
        self._ensure()
        self.lmp.run(1)
        self.steps += 1
        return self.q()

    def q(self):
        return self.lmp.system.get("q")[: self.n].copy()


class Server:
    def __init__(self, lanes: int = 2, hbm_budget: float = 200e9, max_batch_atoms: int = 400_000,
                 ffield: str = DEFAULT_FFIELD, tol: float = 1e-6, mixed: bool = False, batches=None,
                 resident: str = "batched", dt: float = 0.25):
        assert resident in ("batched", "device")
        self.lanes = lanes
        self.batches = batches or [None] * lanes   # reusable per-lane Batch objects (device buffers kept)
        self.budget = hbm_budget
        self.max_batch_atoms = max_batch_atoms
        self.ffield, self.tol, self.mixed = ffield, tol, mixed
        self.resident_mode, self.dt = resident, dt
        self.pending: list[Request] = []
        self.cv = threading.Condition()
        self.in_use = 0.0                         # admitted bytes (batches in flight + device-mode residents)
        self.residents: dict = {}                 # id -> _ResidentState ("batched") or _Resident ("device")
        self.finished: list[Request] = []
        self.closed = False
        self.t0 = 0.0
        self.errors: list[BaseException] = []
        # seconds the workers spent per phase (summed over lanes)
        self.busy = {"pack_load": 0.0, "evaluate": 0.0, "results": 0.0, "resident_steps": 0.0}
        self.nsteps_first = 0

    # ---- admission
    def _take(self, lane):
        """Requests for a lane.  Batched residents: the queued requests in
        arrival order that fit the budget and the batch size cap, new systems
        and steps alike (a step whose system is still being evaluated waits
        for the next batch).  Device residents: all queued steps, else a
        batch of new systems."""
        with self.cv:
            while True:
                if self.pending:
                    if self.resident_mode == "device":
                        steps = [r for r in self.pending if r.kind == "step" and r.resident in self.residents]
                        if steps:
                            for r in steps:
                                self.pending.remove(r)
                            return "step", steps, 0.0
                    take, atoms, nbytes, seen = [], 0, 0.0, set()
                    for r in self.pending:
                        if r.kind == "step":
                            if self.resident_mode == "device":
                                continue
                            st = self.residents.get(r.resident)
                            if st is None or st.busy or r.resident in seen:
                                continue
                            n, b = st.s.n, footprint(st.s)
                        else:
                            n, b = r.system.n, footprint(r.system)
                        if take and (atoms + n > self.max_batch_atoms or self.in_use + nbytes + b > self.budget):
                            break
                        if not take and self.in_use + b > self.budget:
                            break
                        if r.kind == "step":
                            seen.add(r.resident)
                            st.busy = True
                        take.append(r)
                        atoms += n
                        nbytes += b
                    if take:
                        for r in take:
                            self.pending.remove(r)
                        self.in_use += nbytes
                        return "batch", take, nbytes
                if self.closed and not self.pending:
                    return None, [], 0.0
                self.cv.wait(0.002)

    def _release(self, nbytes):
        with self.cv:
            self.in_use -= nbytes
            self.cv.notify_all()

    # ---- workers
    def _worker(self, lane):
        bt = self.batches[lane] or Batch(ffield=self.ffield, elements=ELEMENTS, tol=self.tol, mixed=self.mixed)
        self.batches[lane] = bt
        try:
            while True:
                kind, reqs, nbytes = self._take(lane)
                if kind is None:
                    return
                if kind == "step":
                    self._device_steps(lane, reqs)
                    continue
                self._batch(lane, bt, reqs, nbytes)
        except BaseException as e:  # surface worker failures to run()
            self.errors.append(e)
            with self.cv:
                self.closed = True
                self.cv.notify_all()

    def _device_steps(self, lane, reqs):
        for r in reqs:
            t = time.perf_counter()
            res = self.residents[r.resident]
            with res.lock:
                if res.lmp is None:
                    self.nsteps_first += 1
                r.q = res.step()
                r.step_index = res.steps
            self._finish(r, lane, 1)
            self._account("resident_steps", t)

    def _batch(self, lane, bt, reqs, nbytes):
        """One batch: new systems and (batched mode) resident steps, the
        latter advanced one step and warm-started from their stores."""
        batched = self.resident_mode == "batched"
        t = time.perf_counter()
        syss, hists = [], []
        for r in reqs:
            if r.kind == "step":
                st = self.residents[r.resident]
                r.system = nve_initial(st.s, st.f, self.dt)
                hists.append(st.store)
            else:
                hists.append(None)
            syss.append(r.system)
        bt.load(syss, hist=hists if batched else None)
        t = self._account("pack_load", t)
        bt.evaluate(nonbonded=True, bonded=True)
        # warm starts for the residents' next steps, enqueued before the sync
        stores = {}
        if batched:
            for k, r in enumerate(reqs):
                if r.kind == "step":
                    bt.save_history(k, self.residents[r.resident].store)
                elif r.resident >= 0:
                    stores[k] = bt.save_history(k, fill=True)
        bt.sync()
        t = self._account("evaluate", t)
        qs = bt.charges()
        its = bt.iterations()
        es = bt.energies() + bt.bonded_energies().sum(axis=1)
        hists = bt.histories() if not batched and any(r.resident >= 0 for r in reqs) else None
        grads = bt.gradients() if batched and any(r.resident >= 0 for r in reqs) else None
        for k, (r, q, it) in enumerate(zip(reqs, qs, its)):
            r.q = q.copy()
            r.energy = float(es[k])
            r.iters = tuple(int(v) for v in it)
            if grads is not None and r.resident >= 0:
                r.f = -grads[k]
        with self.cv:
            for k, r in enumerate(reqs):
                if r.kind == "step":
                    st = self.residents[r.resident]
                    st.s = nve_final(r.system, r.f, self.dt)
                    st.f = r.f
                    st.steps += 1
                    r.step_index = st.steps
                    st.busy = False
                elif r.resident >= 0:
                    if batched:
                        self.residents[r.resident] = _ResidentState(r.system, r.f, stores[k])
                    else:
                        r.hist = hists[k]
                        res = _Resident(r.system, self.ffield, self.tol, self.mixed, hist=r.hist)
                        self.residents[r.resident] = res
                        self.in_use += res.bytes
        for r in reqs:
            self._finish(r, lane, len(reqs))
        self._account("results", t)
        self._release(nbytes)

    def _account(self, phase, t):
        now = time.perf_counter()
        with self.cv:
            self.busy[phase] += now - t
        return now

    def _finish(self, r, lane, bsize):
        r.done = time.perf_counter() - self.t0
        r.lane = lane
        r.batch_size = bsize
        with self.cv:
            self.finished.append(r)

    # ---- driver
    def run(self, requests: list[Request]):
        """Feeds the requests at their arrival times; returns them when all are done."""
        self.t0 = time.perf_counter()
        workers = [threading.Thread(target=self._worker, args=(k,), daemon=True) for k in range(self.lanes)]
        for w in workers:
            w.start()
        for r in sorted(requests, key=lambda r: r.arrival):
            wait = r.arrival - (time.perf_counter() - self.t0)
            if wait > 0:
                time.sleep(wait)
            with self.cv:
                self.pending.append(r)
                self.cv.notify_all()
        with self.cv:
            self.closed = True
            self.cv.notify_all()
        for w in workers:
            w.join()
        if self.errors:
            raise self.errors[0]
        return requests


def warm_lanes(batches, max_batch_atoms: int, seed: int = 7):
    """Evaluates one batch of about max_batch_atoms local atoms (periodic
    20,000-atom water boxes, the largest requests) on each lane's Batch, so
    its device buffers start at the size of the largest batch the server
    admits: growing a buffer frees the old one, which waits for the whole
    device and so stalls every lane."""
    big = [make_system("water", 20000, seed=seed + k) for k in range(max(1, max_batch_atoms // 20000))]
    for bt in batches:
        bt.load(big)
        bt.evaluate(nonbonded=True, bonded=True)
        bt.sync()


def summary(reqs: list[Request], wall: float):
    lat = np.array([r.latency for r in reqs if r.q is not None])
    return {
        "requests": len(reqs), "completed": int(lat.size),
        "systems_per_s": round(lat.size / wall, 2),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 2),
                       "p99": round(float(np.percentile(lat, 99)) * 1e3, 2),
                       "max": round(float(lat.max()) * 1e3, 2)},
        "mean_batch": round(float(np.mean([r.batch_size for r in reqs if r.kind == "new"] or [0])), 2),
        "steps": sum(r.kind == "step" for r in reqs),
        "atoms": int(sum(r.system.n for r in reqs if r.system is not None)),
    }
