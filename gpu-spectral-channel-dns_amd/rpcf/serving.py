"""Serving: a persistent request loop over groups (include/rpcf_group.h) with continuous
admission.

Requests (one simulation each: grid, physics, initial-field seed, step count) arrive at any
time.  A request waits in the queue until a slot of its grid shape is free; slots belong to
groups of that shape, and a new group is created only when the free HBM reported by the
runtime covers it (admission by free memory).  Every group has a worker thread that advances
all its running members together by chunks of grouped launches (replayed as a captured HIP
graph while the member list is unchanged) and retires the finished ones.  Newcomers are built
off the running groups' streams: a staging thread per grid shape reserves a free slot, builds
the request's operators and inverses, sets its state and takes its AB2 bootstrap in a
one-slot staging group with its own stream, and the group worker adopts it between two
chunks (device-to-device copies, rpcf_member_adopt), so running members wait only for that
copy.  Groups of different shapes run concurrently on their own streams.

Per request the loop records the queueing delay (arrival to admission, i.e. to the start of
its set-up), the time per 100 steps with its set-up (admission to completion, over its step
count) and without it (joining its group to completion, over the steps after the
bootstrap); a request's result is the same as the simulation run alone.  Requests still
queued or running at the horizon are reported, with lower bounds of their times.  No
counterpart in the reference, which runs one simulation per process.
"""
import ctypes
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import binding as B
from .ensemble import grid_shape

# request grids, Nx x Ny x Nz (streamwise modes x Chebyshev points x spanwise modes)
SERVING_GRIDS = [(32, 33, 32), (64, 65, 64), (128, 129, 128), (256, 129, 256)]


@dataclass
class Request:
    rid: int
    arrival: float  # seconds from the start of the session
    prm: B.Params
    steps: int
    seed: int
    ic: tuple = (1e-2, 4, 8, 0.0)
    keep_state: bool = False
    # filled in by the server
    start: float = None    # admitted: set-up starts
    adopted: float = None  # joined its group: steps after the bootstrap start
    end: float = None
    done: int = 0
    state: tuple = None
    error: str = None

    @property
    def shape(self):
        return (self.prm.mx, self.prm.my, self.prm.mz)


def poisson_requests(rate=20.0, duration=120.0, seed=0, grids=SERVING_GRIDS, steps=(100, 2000)):
    """Poisson arrivals; each request draws a grid, Re ~ U[500, 5000], Ro ~ U[0, 0.5], plane
    Couette or channel by a fair coin, a step count ~ U[steps], and the seeded initial field
    (amplitude 1e-2, |k| <= 4, degree <= 8)"""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= duration:
            return out
        g = grids[rng.integers(len(grids))]
        couette = bool(rng.integers(2))
        prm = B.make_params(**grid_shape(*g), n_pi_x=2.0, n_pi_y=1.0, Re=float(rng.uniform(500.0, 5000.0)),
                            Ro=float(rng.uniform(0.0, 0.5)), dt=0.002,
                            u_profile=(0.0, 1.0, 0.0) if couette else (1.0, 0.0, -1.0),
                            flow_control=0 if couette else 1)
        out.append(Request(len(out), t, prm, int(rng.integers(steps[0], steps[1] + 1)), seed=len(out)))


def free_device_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    if hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) != 0:
        raise B.RpcfError("hipMemGetInfo failed")
    return free.value, total.value


class _Group:
    def __init__(self, server, shape_prm, capacity):
        self.server = server
        self.g = B.Group(server.lib, shape_prm, capacity)
        self.g.graphs(True)
        self.shape = (shape_prm.mx, shape_prm.my, shape_prm.mz)
        self.slots = [None] * capacity  # running request per slot
        self.reserved = set()           # slots promised to a newcomer being staged
        self.ready = []                 # (slot, request, stager) staged and waiting for adoption
        self.thread = threading.Thread(target=self.loop, daemon=True)

    def free_slots(self):
        return [i for i, r in enumerate(self.slots) if r is None and i not in self.reserved]

    def loop(self):
        srv = self.server
        while not srv.stop.is_set():
            with srv.lock:
                ready, self.ready = self.ready, []
            try:
                for slot, req, stager in ready:
                    self.g.adopt(slot, stager.g, 0)
                    req.adopted = srv.now()
                    with srv.lock:
                        self.reserved.discard(slot)
                        self.slots[slot] = req
                    stager.adopted.set()
                    self.retire_if_done(slot)
                active = [i for i, r in enumerate(self.slots) if r is not None]
                if not active:
                    srv.wait_for_work()
                    continue
                chunk = min(srv.chunk, min(self.slots[i].steps - self.slots[i].done for i in active))
                self.g.step(active, chunk)
                for i in active:
                    self.slots[i].done += chunk
                    self.retire_if_done(i)
            except Exception as e:  # a failed group stops; its requests carry the error
                with srv.lock:
                    for r in self.slots:
                        if r is not None:
                            r.error = str(e)
                            srv.finished.append(r)
                self.slots = [None] * len(self.slots)
                srv.record_error(e)
                return

    def retire_if_done(self, slot):
        req = self.slots[slot]
        if req.done < req.steps:
            return
        req.end = self.server.now()
        if req.keep_state:
            req.state = self.g.get_state(slot)
        self.g.release(slot)
        with self.server.lock:
            self.slots[slot] = None
            self.server.finished.append(req)


class _Stager:
    """Set-up of the newcomers of one grid shape in a one-slot staging group (own stream):
    reserve a free slot of a running group, build the request's operators and inverses, set
    its state, take its AB2 bootstrap, hand it to the group and wait until it is adopted."""

    def __init__(self, server, shape_prm):
        self.server = server
        self.g = B.Group(server.lib, shape_prm, 1)
        self.shape = (shape_prm.mx, shape_prm.my, shape_prm.mz)
        self.adopted = threading.Event()
        self.thread = threading.Thread(target=self.loop, daemon=True)

    def loop(self):
        srv = self.server
        dll = B.load(srv.lib)
        while not srv.stop.is_set():
            req = None
            with srv.lock:
                for grp in srv.groups.get(self.shape, []):
                    free = grp.free_slots()
                    if free:
                        req = srv.take(self.shape)
                        if req is not None:
                            slot = free[0]
                            grp.reserved.add(slot)
                            req.start = srv.now()
                        break
            if req is None:
                srv.wait_for_work()
                continue
            try:
                self.g.init(0, req.prm)
                self.g.set_state(0, *B.seed_state(dll, req.prm, req.seed, *req.ic))
                self.g.bootstrap([0])
                req.done = 1
            except Exception as e:
                with srv.lock:
                    grp.reserved.discard(slot)
                    req.error = str(e)
                    srv.finished.append(req)
                srv.record_error(e)
                return
            self.adopted.clear()
            with srv.cv:
                grp.ready.append((slot, req, self))
                srv.cv.notify_all()
            while not self.adopted.wait(0.005):
                if srv.stop.is_set():
                    return


class Server:
    """Continuous-batching request loop on one GPU.

    capacity(shape, slot_bytes) -> slots per new group of that shape; a group is created
    when its bytes (and, for the first group of a shape, its one-slot staging group) fit in
    the free HBM minus `reserve` bytes."""

    def __init__(self, lib=B.HIP_LIB, chunk=10, reserve=8 << 30, capacity=None, max_groups_per_shape=2):
        self.lib = lib
        self.chunk = chunk
        self.reserve = reserve
        self.capacity = capacity or (lambda shape, nbytes: int(max(1, min(16, (24 << 30) // max(nbytes, 1)))))
        self.max_groups = max_groups_per_shape
        self.lock = threading.Lock()
        self.cv = threading.Condition(self.lock)
        self.stop = threading.Event()
        self.queue = []
        self.finished = []
        self.groups = {}
        self.stagers = {}
        self.errors = []
        self.t0 = None

    def now(self):
        return time.perf_counter() - self.t0

    def take(self, shape):  # under the lock
        for k, r in enumerate(self.queue):
            if r.shape == shape:
                return self.queue.pop(k)
        return None

    def wait_for_work(self):
        with self.cv:
            self.cv.wait(timeout=0.002)

    def record_error(self, e):
        with self.lock:
            self.errors.append(e)

    def _maybe_grow(self):
        """create a group (and the shape's stager) for a waiting shape that has no free slot,
        if HBM allows"""
        with self.lock:
            waiting = {r.shape: r for r in self.queue}
            full = {shape: not any(g.free_slots() for g in self.groups.get(shape, [])) for shape in waiting}
        for shape, r in waiting.items():
            groups = self.groups.get(shape, [])
            if not full[shape] or len(groups) >= self.max_groups:
                continue
            nbytes = B.slot_bytes(B.load(self.lib), r.prm)
            extra = 0 if shape in self.stagers else 1
            cap = self.capacity(shape, nbytes)
            free, _ = free_device_bytes()
            while cap > 1 and (cap + extra) * nbytes > free - self.reserve:
                cap -= 1
            if (cap + extra) * nbytes > free - self.reserve:
                continue  # not admitted yet: wait for memory
            g = _Group(self, r.prm, cap)
            with self.lock:
                self.groups.setdefault(shape, []).append(g)
            g.thread.start()
            if extra:
                st = _Stager(self, r.prm)
                self.stagers[shape] = st
                st.thread.start()

    def run(self, requests, horizon):
        """Serve `requests` (arrival times in seconds) until all are finished or `horizon`
        seconds have passed; returns the finished requests (the others keep their state of
        progress: start / adopted / done)."""
        pending = sorted(requests, key=lambda r: r.arrival)
        self.t0 = time.perf_counter()
        k = 0
        while True:
            t = self.now()
            with self.lock:
                nfin = len(self.finished)
            if t >= horizon or (k == len(pending) and nfin >= len(pending)):
                break
            with self.cv:
                while k < len(pending) and pending[k].arrival <= t:
                    self.queue.append(pending[k])
                    k += 1
                self.cv.notify_all()
            self._maybe_grow()
            if self.errors:
                break
            time.sleep(0.001)
        self.wall = self.now()
        self.stop.set()
        for st in self.stagers.values():
            st.thread.join()
        for gs in self.groups.values():
            for g in gs:
                g.thread.join()
                g.g.close()
        for st in self.stagers.values():
            st.g.close()
        if self.errors:
            raise self.errors[0]
        return self.finished


def summarize(finished, submitted, wall):
    """Per request: time per 100 steps with its set-up (admission to completion, over its
    steps) and without it (joining its group to completion, over the steps after the
    bootstrap), queueing delay; p50/p99 over the finished requests, and the same
    percentiles with the requests unfinished at `wall` counted at lower bounds of their
    times (running: (wall - admission) per its steps; never admitted: wall - arrival as
    queueing delay); aggregate grid-point-steps/s of all steps taken over `wall`."""
    ok = [r for r in finished if r.error is None and r.end is not None]
    fin = {r.rid for r in ok}
    running = [r for r in submitted if r.rid not in fin and r.start is not None]
    queued = [r for r in submitted if r.rid not in fin and r.start is None and r.arrival <= wall]
    dll = B.load(B.HIP_LIB)
    pts = {}

    def points(r):
        if r.shape not in pts:
            s = B.sizes(dll, r.prm)
            pts[r.shape] = s["nx"] * s["nz"] * s["ny"]
        return pts[r.shape]

    def pct(a, q):
        return float(np.percentile(a, q)) if len(a) else float("nan")

    per100 = [(r.end - r.start) / r.steps * 100.0 for r in ok]
    per100_steps = [(r.end - r.adopted) / (r.steps - 1) * 100.0 for r in ok if r.steps > 1 and r.adopted is not None]
    qdelay = [r.start - r.arrival for r in ok]
    per100_lb = per100 + [(wall - r.start) / r.steps * 100.0 for r in running]
    qdelay_lb = qdelay + [r.start - r.arrival for r in running] + [wall - r.arrival for r in queued]
    gps = (sum(points(r) * r.done for r in ok) + sum(points(r) * r.done for r in running)) / wall
    return {
        "submitted": len(submitted), "finished": len(ok), "unfinished_running": len(running),
        "unfinished_queued": len(queued),
        "p50_s_per_100_steps": pct(per100, 50), "p99_s_per_100_steps": pct(per100, 99),
        "p50_s_per_100_steps_without_setup": pct(per100_steps, 50),
        "p99_s_per_100_steps_without_setup": pct(per100_steps, 99),
        "p50_queue_s": pct(qdelay, 50), "p99_queue_s": pct(qdelay, 99),
        "p99_s_per_100_steps_with_unfinished_lower_bound": pct(per100_lb, 99),
        "p99_queue_s_with_unfinished_lower_bound": pct(qdelay_lb, 99),
        "grid_point_steps_per_s": float(gps),
    }
