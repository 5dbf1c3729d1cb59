"""Serving a stream of MPM simulations from one GPU: a slot pool in one batch handle,
admission by device footprint, continuous batching.

Every slot of the pool is a simulation of the batch handle with a capacity (grid N^3,
points per body, one contact); a request occupies a free slot (``mpm_gpu_load_sim`` +
uploads), all occupied slots advance together in chunks of steps (``mpm_gpu_step_batch``,
each slot with its own dt), and a slot whose request has run its steps is downloaded and
freed while the others keep going. Requests that arrive while others are on the GPU wait
only for a free slot. The number of slots is what fits a device-memory budget
(``mpm_gpu_footprint`` of one slot).
"""
from __future__ import annotations

import collections
import dataclasses
import time
from typing import List, Optional, Sequence, Tuple

import numpy as np

from mpm_3d_amd import workloads as W
from mpm_3d_amd.gpu import MPM_OPT_GRAPH, MPM_PREC_FP32, GpuSim, load_library, _sim_desc


@dataclasses.dataclass
class CapBody:
    """Capacity of a body slot (what a descriptor needs: a count and material flags)."""
    n: int
    E: Optional[float] = 1e6
    nu: Optional[float] = 0.3
    boundary: bool = True


def slot_workload(dim, N_cap, body_caps, contact=True):
    bodies = [CapBody(int(c)) for c in body_caps]
    contacts = [W.Contact(0, 1, 0.4)] if contact and len(bodies) > 1 else []
    g = (0.0,) * (dim - 1) + (-9.81,)
    return W.Workload("slot", dim, (1.0,) * dim, (N_cap,) * dim, bodies, contacts, 1e-5, 0, g)


def slots_for_budget(dim, N_cap, body_caps, budget_bytes, precision=MPM_PREC_FP32):
    """How many slots of this capacity fit the byte budget (device footprint)."""
    lib = load_library()
    keep = []
    d = _sim_desc(slot_workload(dim, N_cap, body_caps), keep)
    per = lib.mpm_gpu_footprint(d, 1, precision)
    if per <= 0:
        raise ValueError("footprint query failed")
    return max(1, int(budget_bytes // per)), per


class SlotPool:
    def __init__(self, n_slots, dim=3, N_cap=128, body_caps=(2_000_000, 1_000_000), precision=MPM_PREC_FP32,
                 graphs=True):
        self.n_slots = n_slots
        caps = [slot_workload(dim, N_cap, body_caps) for _ in range(n_slots)]
        self.sim = GpuSim(caps, precision=precision, upload=False)
        if graphs:
            self.sim.set_option(MPM_OPT_GRAPH, 1)
        self.req = [None] * n_slots   # request index per slot
        self.loaded = [None] * n_slots  # workload per slot
        self.left = [0] * n_slots     # steps still to run
        self.dt = [1e-5] * n_slots

    def free_slots(self):
        return [s for s in range(self.n_slots) if self.req[s] is None]

    def load(self, s, r, w, steps):
        self.sim.load_sim(s, w)
        self.req[s], self.left[s], self.dt[s], self.loaded[s] = r, steps, w.dt, w

    def result(self, s):
        b0 = self.sim.body0[s]
        return [self.sim.points(b0 + b) for b in range(len(self.loaded[s].bodies))]

    def free(self, s):
        self.sim.clear_sim(s)
        for b in range(len(self.sim.workloads[s].bodies)):
            self.sim.current[self.sim.body0[s] + b] = self.sim.workloads[s].bodies[b]
        self.req[s] = None
        self.loaded[s] = None
        self.left[s] = 0

    def close(self):
        self.sim.close()


def run_stream(pool: SlotPool, requests: Sequence[Tuple[W.Workload, int]], arrivals: Sequence[float],
               chunk_max=64, keep_results=False, log=None):
    """Serve `requests` arriving at `arrivals` (seconds from start, sorted). Returns per-
    request (arrival, start, finish) times, the results if kept, and the makespan."""
    n = len(requests)
    t0 = time.perf_counter()
    queue = collections.deque()
    nxt = 0
    start = [None] * n
    finish = [None] * n
    results = [None] * n
    done = 0
    t_step = t_admin = 0.0
    chunks = 0
    while done < n:
        now = time.perf_counter() - t0
        while nxt < n and arrivals[nxt] <= now:
            queue.append(nxt)
            nxt += 1
        ta = time.perf_counter()
        for s in pool.free_slots():
            if not queue:
                break
            r = queue.popleft()
            w, steps = requests[r]
            pool.load(s, r, w, steps)
            start[r] = time.perf_counter() - t0
        t_admin += time.perf_counter() - ta
        active = [s for s in range(pool.n_slots) if pool.req[s] is not None]
        if not active:
            if nxt < n:
                time.sleep(max(0.0, arrivals[nxt] - (time.perf_counter() - t0)))
            continue
        k = min(chunk_max, min(pool.left[s] for s in active))
        ts = time.perf_counter()
        pool.sim.step_batch(pool.dt, k)
        t_step += time.perf_counter() - ts
        chunks += 1
        ta = time.perf_counter()
        for s in active:
            pool.left[s] -= k
            if pool.left[s] == 0:
                r = pool.req[s]
                if keep_results:
                    results[r] = pool.result(s)
                finish[r] = time.perf_counter() - t0
                pool.free(s)
                done += 1
                if log:
                    log(f"request {r} done: {requests[r][0].n_points} points x {requests[r][1]} steps, "
                        f"latency {finish[r] - arrivals[r]:.3f} s")
        t_admin += time.perf_counter() - ta
    return {"arrival": list(arrivals), "start": start, "finish": finish, "results": results,
            "makespan": time.perf_counter() - t0, "step_s": t_step, "admin_s": t_admin, "chunks": chunks}


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
