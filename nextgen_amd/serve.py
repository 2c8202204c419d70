"""Serving loop: continuous batching of independent solves on one GPU.

Requests (a periodic jittered-Voronoi solve: mesh size, initial field, Re,
Ma, horizon in steps) arrive over time. The server
  * admits arrived requests first-come first-served while the device memory
    held by admitted solves stays within an HBM budget;
  * prepares admitted requests (device layout of the request's own mesh,
    state, initComplete) on worker threads through cg_instance_create while
    the batch keeps stepping (the layout's host-side build holds no library
    lock, so it overlaps the batch's stepping and graph captures);
  * attaches prepared solves to the running batch between steps, so every
    kernel launch covers all resident solves (cg_batch_step);
  * retires a solve at its horizon (its result read back, its memory freed)
    and records its time to solution = completion - arrival.
Each request brings its own mesh and initial field (built before the clock,
as a request's mesh and data files would be read); its device layout is built
with the request and freed when it retires. Layouts given to the server up front (`layouts`) stay
resident and are shared by every request on that mesh (cg_layout_create /
cg_instance_create_on).
A solve's result does not depend on what else is in the batch (the batch
tests check it bit for bit against the solve run alone).

`sequential` is the reference execution: the same requests one at a time in
arrival order, each set up and run alone through cg_create / cg_step (FP64,
or the server's own precision for comparison), with the same queueing
arithmetic (a request starts when it has arrived and the previous one is
done).
"""
from __future__ import annotations

import collections
import concurrent.futures as cf
import time
from dataclasses import dataclass

import numpy as np

from .batch import Batch, Instance, Layout
from .ensemble import Spec
from .ideal_gas import CG_PREC_FP64, CG_PREC_MIXED


@dataclass
class Request:
    rid: int
    arrival: float   # seconds after the start of the run
    spec: Spec       # spec.steps is the horizon


@dataclass
class Result:
    rid: int
    arrival: float
    start: float
    done: float
    cells: int
    steps: int
    ke: float
    enstrophy: float
    time: float

    @property
    def tts(self):
        return self.done - self.arrival


def requests_poisson(rate, count, sizes=None, seed=7, steps_lo=200, steps_hi=2000, cells_lo=2e4, cells_hi=2e6):
    """`count` requests with exponential inter-arrival times of mean 1/rate.
    Cell counts are log-uniform in [cells_lo, cells_hi]; each request has its
    own jittered mesh (lattice side the cube root of its cell count, its own
    seed), or with `sizes` the nearest side of that catalogue (shared meshes).
    Horizons uniform in [steps_lo, steps_hi]; Re, Ma and the initial field as
    in the ensemble."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    cat = None if sizes is None else np.asarray(sorted(sizes))
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        cells = float(np.exp(rng.uniform(np.log(cells_lo), np.log(cells_hi))))
        if cat is None:
            n = int(round(cells ** (1.0 / 3.0)))
        else:
            n = int(cat[np.argmin(np.abs(np.log(cat.astype(float) ** 3) - np.log(cells)))])
        sp = Spec(n=n, mesh_seed=1000 + n, ic="tgv" if rng.random() < 0.5 else "k4",
                  Re=float(rng.uniform(400.0, 3200.0)), Ma=float(rng.uniform(0.05, 0.3)),
                  ic_seed=int(rng.integers(1, 2**31)), steps=int(rng.integers(steps_lo, steps_hi + 1)))
        if cat is None:
            sp.mesh_seed = 100000 + i
        out.append(Request(i, t, sp))
    return out


class MeshPool:
    """The requests' data: meshes by (lattice side, seed), each built once, and
    initial fields by request spec (a request arrives with both, as a mesh file
    and a data file would be read before it is submitted)."""

    def __init__(self):
        self.meshes = {}
        self.states = {}

    def get(self, sp: Spec):
        key = (sp.n, sp.mesh_seed)
        if key not in self.meshes:
            self.meshes[key] = sp.mesh()
        return self.meshes[key]

    def state(self, sp: Spec):
        key = (sp.n, sp.mesh_seed, sp.ic, sp.Ma, sp.ic_seed)
        if key not in self.states:
            self.states[key] = sp.state(self.get(sp))
        return self.states[key]

    def build(self, specs, threads=4):
        """Build the meshes and initial fields of `specs` (several at a time: the
        builders release the interpreter while they work)."""
        todo = {}
        for sp in specs:
            todo.setdefault((sp.n, sp.mesh_seed), sp)
        todo = {k: sp for k, sp in todo.items() if k not in self.meshes}
        with cf.ThreadPoolExecutor(max_workers=threads) as ex:
            for key, m in zip(todo, ex.map(lambda sp: sp.mesh(), todo.values())):
                self.meshes[key] = m
            list(ex.map(self.state, specs))


def build_layouts(pool: MeshPool, precision):
    """Resident device layouts of every mesh in the pool."""
    return {key: Layout(m, precision) for key, m in pool.meshes.items()}


class Server:
    """`batch_factory(precision)` and `prepare(request)` default to the GPU
    Batch and cg_instance_create (the request's own layout) or
    cg_instance_create_on a resident layout; they are parameters so the
    scheduling logic can be exercised without a GPU. `workers` requests are
    prepared at a time. The HBM budget covers the admitted solves' memory,
    their own layouts and the resident layouts."""

    def __init__(self, pool: MeshPool, precision=CG_PREC_MIXED, hbm_budget=96e9, max_chunk=16,
                 bytes_per_cell=None, log=None, batch_factory=None, prepare=None, layouts=None, workers=3):
        self.pool = pool
        self.precision = precision
        self.budget = float(hbm_budget)
        self.max_chunk = int(max_chunk)
        self.workers = max(1, int(workers))
        # admission estimate of a request's device bytes per cell (its state
        # and work buffers, and its own layout), raised to what requests hold
        self.bpc = bytes_per_cell or (5000.0 if precision == CG_PREC_FP64 else 3000.0)
        self.log = log or (lambda s: None)
        self.batch_factory = batch_factory or Batch
        self.prepare = prepare or self._prepare
        self.layouts = layouts if layouts is not None else {}
        self.peak_bytes = 0.0
        self.prep_seconds = {}  # request id -> seconds its preparation took

    def _timed_prepare(self, req: Request):
        t0 = time.perf_counter()
        inst = self.prepare(req)
        self.prep_seconds[req.rid] = time.perf_counter() - t0
        return inst

    def _prepare(self, req: Request):
        m = self.pool.get(req.spec)
        key = (req.spec.n, req.spec.mesh_seed)
        deck = req.spec.deck(self.precision, check_interval=100)
        st = self.pool.state(req.spec)
        if key in self.layouts:
            return Instance(m, deck, *st, layout=self.layouts[key])
        inst = Instance(m, deck, *st)  # its own layout, freed with it
        inst.private_layout = True
        return inst

    def run(self, requests):
        reqs = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting = collections.deque()
        preparing = {}          # future -> (request, reserved bytes)
        active = {}             # batch id -> [request, steps left, start, bytes, cells]
        results = []
        layout_bytes = {k: float(v.bytes()) for k, v in self.layouts.items()} if self.layouts else {}
        used = sum(layout_bytes.values())
        batch = self.batch_factory(self.precision)
        pool = cf.ThreadPoolExecutor(max_workers=self.workers)
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        steps_run = 0
        cell_steps = 0
        # host-side time accounting of the loop: inside batch steps (the GPU is
        # busy for about this long), attaching, retiring, waiting with nothing
        # resident
        t_step = t_attach = t_retire = t_idle = 0.0
        while reqs or waiting or preparing or active:
            t = now()
            while reqs and reqs[0].arrival <= t:
                waiting.append(reqs.popleft())
            # admission by HBM budget, first come first served
            while waiting:
                r = waiting[0]
                need = self.bpc * r.spec.n ** 3
                if used + need > self.budget and (active or preparing):
                    break
                waiting.popleft()
                used += need
                self.peak_bytes = max(self.peak_bytes, used)
                preparing[pool.submit(self._timed_prepare, r)] = (r, need)
            # attach what is ready
            ta = time.perf_counter()
            for fut in [f for f in preparing if f.done()]:
                r, need = preparing.pop(fut)
                inst = fut.result()
                own, lay = inst.bytes()
                key = (r.spec.n, r.spec.mesh_seed)
                hold = float(own)
                if getattr(inst, "private_layout", False):
                    hold += lay  # the request's own layout, freed when it retires
                elif key not in layout_bytes:  # a shared layout stays resident
                    layout_bytes[key] = float(lay)
                    used += lay
                used += hold - need
                self.peak_bytes = max(self.peak_bytes, used)
                self.bpc = max(self.bpc, hold / r.spec.n ** 3)
                iid = batch.attach(inst)
                active[iid] = [r, r.spec.steps, now(), hold, r.spec.n ** 3]
            t_attach += time.perf_counter() - ta
            if not active:
                # nothing to step: wait for a preparation to finish or the next
                # request to arrive, whichever comes first
                ti = time.perf_counter()
                nxt = [f for f in preparing]
                until = max(0.0, reqs[0].arrival - now()) if reqs else None
                if nxt:
                    cf.wait(nxt, timeout=until, return_when=cf.FIRST_COMPLETED)
                elif reqs:
                    time.sleep(min(until, 0.05))
                t_idle += time.perf_counter() - ti
                continue
            k = min(self.max_chunk, min(a[1] for a in active.values()))
            ts = time.perf_counter()
            batch.step(k)
            t_step += time.perf_counter() - ts
            steps_run += k
            cell_steps += k * sum(a[4] for a in active.values())
            tr = time.perf_counter()
            for iid in list(active):
                a = active[iid]
                a[1] -= k
                if a[1] == 0:
                    d = batch.diag(iid)
                    r = a[0]
                    if d.step < r.spec.steps:
                        # a NaN rollback took the solve back: the horizon counts net
                        # steps, as the reference run loop does
                        a[1] = r.spec.steps - d.step
                        continue
                    results.append(Result(r.rid, r.arrival, a[2], now(), a[4], r.spec.steps, d.ke, d.enstrophy, d.time))
                    batch.remove(iid)
                    used -= a[3]
                    del active[iid]
            t_retire += time.perf_counter() - tr
        makespan = now()
        pool.shutdown()
        batch.close()
        return results, {"makespan": makespan, "batch_steps": steps_run, "cell_steps": cell_steps,
                         "host_seconds": {"stepping": t_step, "attaching": t_attach, "retiring": t_retire,
                                          "idle_nothing_resident": t_idle}}


def sequential(pool: MeshPool, requests, precision=CG_PREC_FP64, layouts=None):
    """One request at a time in arrival order, each alone on the GPU: set up
    (cg_instance_create: its own layout, or the mesh's resident layout when
    `layouts` has it), then stepped alone with graph replays (a batch of one).
    Service times are measured; waiting follows from the arrival times."""
    results = []
    clock = 0.0
    for r in sorted(requests, key=lambda q: q.arrival):
        m = pool.get(r.spec)
        key = (r.spec.n, r.spec.mesh_seed)
        t0 = time.perf_counter()
        lay = layouts.get(key) if layouts else None
        b = Batch(precision)
        iid = b.attach(Instance(m, r.spec.deck(precision, check_interval=100), *pool.state(r.spec), layout=lay))
        b.step(r.spec.steps)
        d = b.diag(iid)
        b.close()
        service = time.perf_counter() - t0
        start = max(clock, r.arrival)
        clock = start + service
        results.append(Result(r.rid, r.arrival, start, clock, r.spec.n ** 3, r.spec.steps, d.ke, d.enstrophy, d.time))
    return results


def summarize(results, makespan=None):
    tts = np.array([r.tts for r in results])
    # both executions measure time from the same origin (the arrival clock's zero)
    span = makespan if makespan is not None else max(r.done for r in results)
    return {"solves": len(results), "makespan_s": span, "solves_per_s": len(results) / span,
            "tts_p50_s": float(np.percentile(tts, 50)), "tts_p99_s": float(np.percentile(tts, 99)),
            "tts_mean_s": float(tts.mean())}
