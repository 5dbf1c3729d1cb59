"""Serving a stream of MPM simulations from one GPU: a slot pool in one batch handle,
admission by device footprint, continuous batching.

Every slot of the pool is a simulation of the batch handle with a capacity (grid N^3,
points per body, one contact); slots come in classes of different capacity, so that the
device memory a slot holds follows the requests it is meant for. A request occupies the
smallest free slot that fits it (``mpm_gpu_load_sim`` + uploads), all occupied slots
advance together in chunks of steps (``mpm_gpu_step_batch``, each slot with its own dt),
and a slot whose request has run its steps is downloaded and freed while the others keep
going. Requests that arrive while others are on the GPU wait only for a free slot that
fits them. The reference has no serving layer: each request stands for a Job of its own
(config.cpp:176-213, job.hpp:76-105) stepped by ExplicitUSL::step (explicit_usl.cpp:48-87).

Failure handling, per request (the others are unaffected):
  refused    no slot of the pool can ever hold it (grid, bodies or points above every
             slot's capacity); it is never loaded
  cancelled  ``cancel_at`` reached while it waited (dropped from the queue) or ran (its
             slot is freed at the next chunk boundary)
  failed     its state at completion holds a non-finite value
  done       otherwise

``run_stream`` drives any executor with the methods of ``SlotPool`` (n_slots, fits,
fits_any, capacity, free_slots, load, active_slots, remaining, request_of, step, result,
free), so the admission logic is tested on the CPU against a mock executor.
"""
from __future__ import annotations

import dataclasses
import time
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from mpm_3d_amd import workloads as W
from mpm_3d_amd.gpu import (MPM_OPT_AUTOTUNE, MPM_OPT_GRAPH, MPM_OPT_KEEP_L, MPM_OPT_SORT_INTERVAL, MPM_PREC_FP32,
                             GpuSim, load_library, _sim_desc)


@dataclasses.dataclass
class CapBody:
    """Capacity of a body slot (what a descriptor needs: a count and material flags)."""
    n: int
    E: Optional[float] = 1e6
    nu: Optional[float] = 0.3
    boundary: bool = True


@dataclasses.dataclass
class SlotClass:
    """`count` slots of grid N_cap^dim holding up to body_caps[b] points in body b."""
    count: int
    N_cap: int
    body_caps: Tuple[int, ...]


def slot_workload(dim, N_cap, body_caps, contact=True):
    bodies = [CapBody(int(c)) for c in body_caps]
    contacts = [W.Contact(0, 1, 0.4)] if contact and len(bodies) > 1 else []
    g = (0.0,) * (dim - 1) + (-9.81,)
    return W.Workload("slot", dim, (1.0,) * dim, (N_cap,) * dim, bodies, contacts, 1e-5, 0, g)


def slot_footprint(dim, N_cap, body_caps, precision=MPM_PREC_FP32):
    """Device bytes of one slot of this capacity (mpm_gpu_footprint)."""
    lib = load_library()
    keep = []
    d = _sim_desc(slot_workload(dim, N_cap, body_caps), keep)
    per = lib.mpm_gpu_footprint(d, 1, precision)
    if per <= 0:
        raise ValueError("footprint query failed")
    return int(per)


def slots_for_budget(dim, N_cap, body_caps, budget_bytes, precision=MPM_PREC_FP32):
    """How many slots of this capacity fit the byte budget (device footprint)."""
    per = slot_footprint(dim, N_cap, body_caps, precision)
    return max(1, int(budget_bytes // per)), per


def request_fits(w, N_cap, body_caps, dim=3):
    """Whether workload w fits a slot of grid N_cap and these body capacities."""
    if w.dim != dim or any(n > N_cap for n in w.N) or len(w.bodies) > len(body_caps):
        return False
    if len(w.contacts) > (1 if len(body_caps) > 1 else 0):
        return False
    return all(b.n <= c for b, c in zip(w.bodies, body_caps))


def classes_for_budget(dim, class_shapes, budget_bytes, precision=MPM_PREC_FP32, shares=None):
    """Slot classes sized by device footprint: class_shapes is a list of (N_cap, body_caps);
    the byte budget is split between them by `shares` (default: equal), each class holding
    as many slots as its share's bytes pay for (at least one)."""
    shares = shares or [1.0] * len(class_shapes)
    tot = float(sum(shares))
    out, per = [], []
    for (N_cap, caps), sh in zip(class_shapes, shares):
        f = slot_footprint(dim, N_cap, caps, precision)
        out.append(SlotClass(max(1, int(budget_bytes * sh / tot // f)), N_cap, tuple(caps)))
        per.append(f)
    return out, per


class SlotPool:
    """Slots in one batch handle. Either `n_slots` slots of one capacity (N_cap, body_caps)
    or `classes`, a list of SlotClass. keep_L=False: no velocity-gradient plane (72 B per
    point slot), which only mpm_gpu_download_aux reads. sort_interval: steps between
    re-sorts of the pool's points (32, the value the autotuner picks for the 16M column and
    the 512-simulation ensemble; a pool's mix changes as it runs, so it is not tuned here)."""

    def __init__(self, n_slots=None, dim=3, N_cap=128, body_caps=(2_000_000, 1_000_000), precision=MPM_PREC_FP32,
                 graphs=True, classes=None, keep_L=False, sort_interval=32, autotune=False):
        if classes is None:
            classes = [SlotClass(int(n_slots), N_cap, tuple(body_caps))]
        self.dim = dim
        self.classes = list(classes)
        self.cap = []  # (N_cap, body_caps) per slot
        caps = []
        for c in self.classes:
            for _ in range(c.count):
                caps.append(slot_workload(dim, c.N_cap, c.body_caps))
                self.cap.append((c.N_cap, tuple(c.body_caps)))
        self.n_slots = len(caps)
        self.footprint = [slot_footprint(dim, n, b, precision) for n, b in self.cap]
        self.sim = GpuSim(caps, precision=precision, upload=False)
        if not keep_L:
            self.sim.set_option(MPM_OPT_KEEP_L, 0)
        self.sim.set_option(MPM_OPT_SORT_INTERVAL, int(sort_interval))
        if autotune:  # launch shapes, on the first chunk long enough (28 steps)
            self.sim.set_option(MPM_OPT_AUTOTUNE, 1)
        if graphs:
            self.sim.set_option(MPM_OPT_GRAPH, 1)
        self.req = [None] * self.n_slots     # request index per slot
        self.loaded = [None] * self.n_slots  # workload per slot
        self.left = [0] * self.n_slots       # steps still to run
        self.dt = [1e-5] * self.n_slots

    # -- executor interface used by run_stream ------------------------------------
    def fits(self, s, w):
        return request_fits(w, self.cap[s][0], self.cap[s][1], self.dim)

    def fits_any(self, w):
        return any(self.fits(s, w) for s in range(self.n_slots))

    def capacity(self, s):
        """Ordering key of slot s: its device footprint (smaller slots are taken first)."""
        return self.footprint[s]

    def free_slots(self):
        return [s for s in range(self.n_slots) if self.req[s] is None]

    def active_slots(self):
        return [s for s in range(self.n_slots) if self.req[s] is not None]

    def remaining(self, s):
        return self.left[s]

    def request_of(self, s):
        return self.req[s]

    def load(self, s, r, w, steps, prep=None):
        self.sim.load_sim(s, w, preps=prep)
        self.req[s], self.left[s], self.dt[s], self.loaded[s] = r, steps, w.dt, w

    def step(self, k):
        self.sim.step_batch(self.dt, k)
        for s in self.active_slots():
            self.left[s] -= k

    def result(self, s):
        return self.finish(self.fetch(s))

    # -- host work split from the device work, so run_stream can overlap it with stepping --
    def prepare(self, w):
        """Host arrays of w's uploads (no GPU call)."""
        return [GpuSim.upload_arrays(b, self.dim) for b in w.bodies]

    def fetch(self, s):
        """Download slot s's bodies as component planes (GPU calls)."""
        b0 = self.sim.body0[s]
        return [self.sim.points_planes(b0 + b) for b in range(len(self.loaded[s].bodies))]

    def finish(self, fetched):
        """fetch() output in the result layout (no GPU call)."""
        return [GpuSim.rows(p) for p in fetched]

    def free(self, s):
        self.sim.clear_sim(s)
        for b in range(len(self.sim.workloads[s].bodies)):
            self.sim.current[self.sim.body0[s] + b] = self.sim.workloads[s].bodies[b]
        self.req[s] = None
        self.loaded[s] = None
        self.left[s] = 0

    def close(self):
        self.sim.close()


def check_result(w, bodies):
    """None if a finished request's state is finite, else why it failed. (A point whose
    position became NaN leaves the grid and is deactivated; its velocity keeps the NaN.)"""
    for bi, st in enumerate(bodies):
        for key in ("x", "xt", "T6", "v"):
            a = st.get(key)
            if a is not None and not np.all(np.isfinite(a)):
                return f"non-finite {key} in body {bi}"
    return None


def run_stream(pool, requests: Sequence[Tuple[W.Workload, int]], arrivals: Sequence[float], chunk_max=64,
               keep_results=False, log=None, cancel_at: Optional[Dict[int, float]] = None, clock=None,
               sleep=None, overlap: Optional[bool] = None, prep_ahead=4):
    """Serve `requests` arriving at `arrivals` (seconds from start, sorted). Returns per-
    request (arrival, start, finish) times and status, the results if kept, and the
    makespan. Admission is first come, first served over the slots a request fits,
    smallest slot first: a free slot goes to the earliest waiting request that fits it, so
    a large request waiting for a large slot gets it before any later request.

    overlap: each chunk of steps runs on a worker thread while this thread does the host work
    around it: the result layout and non-finite check of the requests that finished at the
    last chunk boundary (downloaded there, their slots already free) and the upload arrays of
    the first `prep_ahead` waiting requests. Only calls into the executor's device (load,
    fetch, free) stay between chunks. Needs an executor with prepare, fetch and finish; by
    default on when it has them."""
    clock = clock or time.perf_counter
    sleep = sleep or time.sleep
    cancel_at = dict(cancel_at or {})
    if overlap is None:
        overlap = all(hasattr(pool, a) for a in ("prepare", "fetch", "finish"))
    n = len(requests)
    t0 = clock()
    queue: List[int] = []
    nxt = 0
    start = [None] * n
    finish = [None] * n
    status = ["pending"] * n
    results = [None] * n
    done = 0
    t_step = t_admin = 0.0
    t_load = t_result = t_free = 0.0  # parts of t_admin
    t_host = 0.0  # host work done while a chunk stepped (overlap)
    chunks = 0
    prepared: Dict[int, list] = {}  # request -> its upload arrays (overlap)
    fetched: List[Tuple[int, list]] = []  # finished requests downloaded, not yet checked

    def close(r, st):
        nonlocal done
        status[r] = st
        finish[r] = clock() - t0
        done += 1

    def complete(r, res):
        why = check_result(requests[r][0], res)
        if keep_results:
            results[r] = res
        close(r, "done" if why is None else f"failed: {why}")
        if log:
            log(f"request {r} {status[r]}: {requests[r][0].n_points} points x {requests[r][1]} steps, "
                f"latency {finish[r] - arrivals[r]:.3f} s")

    def finish_fetched():
        for r, raw in fetched:
            complete(r, pool.finish(raw))
        fetched.clear()

    worker = None
    if overlap:
        from concurrent.futures import ThreadPoolExecutor
        worker = ThreadPoolExecutor(max_workers=1)
    try:
        while done < n:
            now = clock() - t0
            while nxt < n and arrivals[nxt] <= now:
                w = requests[nxt][0]
                if pool.fits_any(w):
                    queue.append(nxt)
                else:
                    close(nxt, "refused")
                    if log:
                        log(f"request {nxt} refused: larger than every slot")
                nxt += 1
            ta = clock()
            # cancellations due: waiting requests leave the queue, running ones free their slot
            for r, tc in list(cancel_at.items()):
                if tc > now or status[r] != "pending":
                    continue
                if r in queue:
                    queue.remove(r)
                    prepared.pop(r, None)
                    close(r, "cancelled")
                    del cancel_at[r]
                else:
                    for s in pool.active_slots():
                        if pool.request_of(s) == r:
                            pool.free(s)
                            close(r, "cancelled")
                            del cancel_at[r]
                            break
            # admission: in queue order, each request into the smallest free slot it fits;
            # one that fits none waits, and later ones may take what is left
            free = sorted(pool.free_slots(), key=pool.capacity)
            for r in list(queue):
                if not free:
                    break
                w, steps = requests[r]
                cand = [s for s in free if pool.fits(s, w)]
                if cand:
                    tl = clock()
                    if overlap:
                        pool.load(cand[0], r, w, steps, prep=prepared.pop(r, None))
                    else:
                        pool.load(cand[0], r, w, steps)
                    t_load += clock() - tl
                    start[r] = clock() - t0
                    queue.remove(r)
                    free.remove(cand[0])
            t_admin += clock() - ta
            active = pool.active_slots()
            if not active:
                if fetched:  # nothing steps: check the last results now
                    ta = clock()
                    finish_fetched()
                    t_result += clock() - ta
                    t_admin += clock() - ta
                    continue
                if nxt < n:
                    wait = arrivals[nxt] - (clock() - t0)
                    nexts = [tc for r, tc in cancel_at.items() if status[r] == "pending"]
                    if nexts:
                        wait = min(wait, min(nexts) - (clock() - t0))
                    if wait > 0:
                        sleep(wait)
                elif queue:
                    raise RuntimeError("requests wait for slots while none are in use")
                continue
            k = min(chunk_max, min(pool.remaining(s) for s in active))
            ts = clock()
            if overlap:
                job = worker.submit(pool.step, k)
                th = clock()
                finish_fetched()
                for r in queue[:prep_ahead]:
                    if r not in prepared:
                        prepared[r] = pool.prepare(requests[r][0])
                t_host += clock() - th
                job.result()
            else:
                pool.step(k)
            t_step += clock() - ts
            chunks += 1
            ta = clock()
            for s in active:
                if pool.remaining(s) == 0:
                    r = pool.request_of(s)
                    tr = clock()
                    if overlap:
                        fetched.append((r, pool.fetch(s)))
                        t_result += clock() - tr
                    else:
                        res = pool.result(s)
                        t_result += clock() - tr
                    tf = clock()
                    pool.free(s)
                    t_free += clock() - tf
                    if not overlap:
                        complete(r, res)
            t_admin += clock() - ta
    finally:
        if worker is not None:
            worker.shutdown(wait=True)
    return {"arrival": list(arrivals), "start": start, "finish": finish, "status": status, "results": results,
            "makespan": clock() - t0, "step_s": t_step, "admin_s": t_admin, "chunks": chunks,
            "load_s": t_load, "result_s": t_result, "free_s": t_free, "host_overlapped_s": t_host,
            "overlap": bool(overlap)}


def poisson_arrivals(n, rate, seed=0xC0FFEE + 5):
    rng = np.random.default_rng(seed)
    return list(np.cumsum(rng.exponential(1.0 / rate, size=n)))


def run_alone(w, steps, precision=MPM_PREC_FP32):
    """One request on its own handle (create, upload, step, download)."""
    sim = GpuSim(w, precision=precision)
    sim.step(w.dt, steps)
    out = [sim.points(b) for b in range(len(w.bodies))]
    sim.close()
    return out
