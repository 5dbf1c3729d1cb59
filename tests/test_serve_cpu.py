"""The serving loop's admission and failure handling against a mock executor (no GPU):
first come, first served over the slots a request fits, smallest slot first, no starvation
of large requests by small ones, refusal of requests no slot can hold, cancellation of
waiting and running requests, and non-finite results reported as failures without
disturbing the others; each with the host work done inline and overlapped with stepping."""
import numpy as np
import pytest

from mpm_3d_amd import serve as S
from mpm_3d_amd import workloads as W


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += dt


class MockPool:
    """Slots with capacities; every step takes `step_s` of mock time; a request whose name
    contains 'nan' yields a NaN velocity."""

    def __init__(self, caps, clock, step_s=0.01):
        self.cap = caps  # (N_cap, body_caps) per slot
        self.n_slots = len(caps)
        self.clock = clock
        self.step_s = step_s
        self.req = [None] * self.n_slots
        self.w = [None] * self.n_slots
        self.left = [0] * self.n_slots
        self.loads = []  # (slot, request, time)

    def fits(self, s, w):
        return S.request_fits(w, self.cap[s][0], self.cap[s][1])

    def fits_any(self, w):
        return any(self.fits(s, w) for s in range(self.n_slots))

    def capacity(self, s):
        N, caps = self.cap[s]
        return N ** 3 * 100 + sum(caps) * 68

    def free_slots(self):
        return [s for s in range(self.n_slots) if self.req[s] is None]

    def active_slots(self):
        return [s for s in range(self.n_slots) if self.req[s] is not None]

    def remaining(self, s):
        return self.left[s]

    def request_of(self, s):
        return self.req[s]

    def load(self, s, r, w, steps):
        assert self.req[s] is None and self.fits(s, w)
        self.req[s], self.w[s], self.left[s] = r, w, steps
        self.loads.append((s, r, self.clock()))

    def step(self, k):
        self.clock.t += k * self.step_s
        for s in self.active_slots():
            self.left[s] -= k
            assert self.left[s] >= 0

    def result(self, s):
        w = self.w[s]
        out = []
        for b in w.bodies:
            xt = np.zeros((b.n, 3))
            if "nan" in w.name:
                xt[0, 0] = np.nan
            out.append({"x": np.zeros((b.n, 3)), "xt": xt, "T6": np.zeros((b.n, 6)), "v": np.ones(b.n),
                        "active": np.ones(b.n, np.int32)})
        return out

    def free(self, s):
        self.req[s] = None
        self.w[s] = None
        self.left[s] = 0


class MockPoolOverlap(MockPool):
    """MockPool with the host/device split of SlotPool (prepare, fetch, finish), so the
    serving loop steps each chunk on a worker thread and does the host work meanwhile."""

    def __init__(self, *a, **k):
        super().__init__(*a, **k)
        self.prepared = []   # requests whose upload arrays were made
        self.with_prep = []  # requests loaded with prepared arrays

    def load(self, s, r, w, steps, prep=None):
        super().load(s, r, w, steps)
        if prep is not None:
            assert prep == ("prep", w.name)
            self.with_prep.append(r)

    def prepare(self, w):
        self.prepared.append(w.name)
        return ("prep", w.name)

    def fetch(self, s):
        return ("fetched", self.result(s))

    def finish(self, fetched):
        assert fetched[0] == "fetched"
        return fetched[1]


def pool_of(overlap, caps, clk):
    return (MockPoolOverlap if overlap else MockPool)(caps, clk)


def req(name, N, npts, nbody=1):
    bodies = [W.Body(f"b{i}", np.ones(npts), np.ones(npts), np.zeros((npts, 3)), np.zeros((npts, 3)),
                     np.ones(npts, np.int32), 1e6, 0.3, True) for i in range(nbody)]
    contacts = [W.Contact(0, 1, 0.4)] if nbody == 2 else []
    return W.Workload(name, 3, (1.0,) * 3, (N,) * 3, bodies, contacts, 1e-5, 100, (0, 0, -9.81))


SMALL = (64, (300, 300))
BIG = (128, (2000, 1000))


@pytest.mark.parametrize("overlap", [False, True])
def test_fifo_smallest_slot_and_reuse(overlap):
    clk = Clock()
    pool = pool_of(overlap, [BIG, SMALL, SMALL], clk)
    reqs = [(req(f"r{i}", 40, 100), 10 * (i + 1)) for i in range(6)]
    out = S.run_stream(pool, reqs, [0.0] * 6, chunk_max=8, clock=clk, sleep=clk.sleep)
    assert out["overlap"] == overlap
    assert out["status"] == ["done"] * 6
    if overlap:  # the three that waited were loaded with upload arrays made during stepping
        assert sorted(pool.with_prep) == [3, 4, 5]
    # the two small slots take the first two; the big slot is used only when they are busy
    first = {r: s for s, r, _ in pool.loads[:3]}
    assert first[0] in (1, 2) and first[1] in (1, 2) and first[2] == 0
    # first come, first served: start times follow arrival order
    assert out["start"] == sorted(out["start"])
    assert len(pool.loads) == 6 and len({s for s, _, _ in pool.loads}) == 3
    for r, (w, st) in enumerate(reqs):
        assert out["finish"][r] >= out["start"][r] + st * pool.step_s - 1e-12


@pytest.mark.parametrize("overlap", [False, True])
def test_large_request_not_starved_by_small_ones(overlap):
    clk = Clock()
    pool = pool_of(overlap, [BIG, SMALL], clk)
    # r0 occupies the big slot; r1 (big) waits for it; r2.. (small) must not take the big
    # slot when it frees, although they could use it
    reqs = [(req("r0", 100, 1500), 20), (req("r1", 100, 1500), 20)]
    reqs += [(req(f"s{i}", 40, 100), 50) for i in range(4)]
    out = S.run_stream(pool, reqs, [0.0] * len(reqs), chunk_max=5, clock=clk, sleep=clk.sleep)
    assert out["status"] == ["done"] * len(reqs)
    big_loads = [r for s, r, _ in pool.loads if s == 0]
    assert big_loads[:2] == [0, 1]


@pytest.mark.parametrize("overlap", [False, True])
def test_refused_cancelled_and_failed(overlap):
    clk = Clock()
    pool = pool_of(overlap, [SMALL, SMALL], clk)
    reqs = [
        (req("ok0", 40, 100), 40),
        (req("too_big_grid", 96, 100), 10),   # N above every slot
        (req("too_many_points", 40, 400), 10),  # points above every slot
        (req("nan1", 40, 100), 20),           # its state turns non-finite
        (req("waits_then_cancelled", 40, 100), 10),
        (req("cancelled_running", 40, 100), 400),
        (req("ok1", 40, 100, nbody=2), 30),
    ]
    arrivals = [0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.05]
    cancel = {4: 0.05, 5: 1.0}
    logs = []
    out = S.run_stream(pool, reqs, arrivals, chunk_max=10, clock=clk, sleep=clk.sleep, cancel_at=cancel,
                       log=logs.append)
    st = out["status"]
    assert st[0] == "done" and st[6] == "done"
    assert st[1] == "refused" and st[2] == "refused"
    assert st[3].startswith("failed: non-finite xt")
    assert st[4] == "cancelled" and out["start"][4] is None
    assert st[5] == "cancelled" and out["start"][5] is not None
    assert out["finish"][5] < out["start"][5] + 400 * pool.step_s
    assert not any(r in (1, 2, 4) for _, r, _ in pool.loads)


@pytest.mark.parametrize("overlap", [False, True])
def test_waits_for_arrivals(overlap):
    clk = Clock()
    pool = pool_of(overlap, [SMALL], clk)
    reqs = [(req(f"r{i}", 40, 100), 10) for i in range(3)]
    arrivals = [0.5, 1.0, 1.01]
    out = S.run_stream(pool, reqs, arrivals, chunk_max=10, clock=clk, sleep=clk.sleep)
    assert out["status"] == ["done"] * 3
    for r in range(3):
        assert out["start"][r] >= arrivals[r]
    assert abs(out["start"][0] - 0.5) < 1e-9 and abs(out["start"][1] - 1.0) < 1e-9
    assert abs(out["start"][2] - 1.1) < 1e-9  # waits for the only slot


def test_request_fits():
    assert S.request_fits(req("a", 64, 300), *SMALL)
    assert not S.request_fits(req("a", 65, 300), *SMALL)
    assert not S.request_fits(req("a", 64, 301), *SMALL)
    assert S.request_fits(req("a", 64, 300, nbody=2), *SMALL)
    assert not S.request_fits(req("a", 64, 10, nbody=2), 64, (300,))
