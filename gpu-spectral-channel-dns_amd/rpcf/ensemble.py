"""Ensembles of independent simulations of uneven grid sizes on one GPU.

Members are grouped by grid shape; each shape is one group (include/rpcf_group.h), so a
step of the whole ensemble is one set of grouped launches per shape, and the shapes run
concurrently on their own streams (one host thread per group; the ctypes calls release
the GIL).  The one-at-a-time baseline runs the same members as single simulations in
sequence.  No counterpart in the reference, which runs one simulation per process
(source/main_loop.cpp:15-115).
"""
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import binding as B

# grids of the mixed-size ensemble, Nx x Ny x Nz = streamwise modes x Chebyshev points x
# spanwise modes, with 3/2 padding: mx = 3/2 Nx, my = 3/2 Nz, mz = 4 (Ny - 1)
ENSEMBLE_GRIDS = [(16, 33, 16), (32, 49, 32), (32, 65, 32), (64, 65, 64)]


def grid_shape(nx, ny_cheb, nz_span):
    return dict(mx=nx * 3 // 2, my=nz_span * 3 // 2, mz=4 * (ny_cheb - 1))


@dataclass
class Member:
    """one simulation: its parameters, initial-field recipe and seed"""
    prm: B.Params
    seed: int
    ic: tuple = (1e-2, 4, 8, 0.0)  # amplitude, |k| <= kmax, Chebyshev degree <= degmax, k-exponent
    tag: str = ""

    @property
    def shape(self):
        return (self.prm.mx, self.prm.my, self.prm.mz)


def mixed_members(n=64, seed=0, dt=0.002, precision=1):
    """the mixed-size ensemble: member i has seed i; grid uniform over ENSEMBLE_GRIDS,
    Re ~ U[500, 3000], Ro ~ U[0, 0.5], plane Couette (U = y) or pressure-driven channel
    (U = 1 - y^2, constant flow rate) by a fair coin; Lx = 2 pi, Lz = pi; precision 1 (the
    mixed form: FP32 transforms, FP64 solves) or 0 (FP64)"""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        g = ENSEMBLE_GRIDS[rng.integers(len(ENSEMBLE_GRIDS))]
        re = float(rng.uniform(500.0, 3000.0))
        ro = float(rng.uniform(0.0, 0.5))
        couette = bool(rng.integers(2))
        prm = B.make_params(**grid_shape(*g), n_pi_x=2.0, n_pi_y=1.0, Re=re, Ro=ro, dt=dt,
                            u_profile=(0.0, 1.0, 0.0) if couette else (1.0, 0.0, -1.0),
                            flow_control=0 if couette else 1, precision=precision)
        out.append(Member(prm, seed=i, tag=f"{g[0]}x{g[1]}x{g[2]} {'couette' if couette else 'channel'}"))
    return out


def initial_state(dll, m):
    return B.seed_state(dll, m.prm, m.seed, *m.ic)


@dataclass
class GroupRun:
    shape: tuple
    members: list  # indices into the member list
    group: object = None
    seconds: float = 0.0
    error: object = None
    states: dict = field(default_factory=dict)


def run_grouped(members, nsteps, lib=B.HIP_LIB, graphs=True, keep_states=False):
    """Bootstrap every member, then time `nsteps` steps of the whole ensemble (one group
    per shape, the groups concurrently).  Returns (wall seconds of the timed steps,
    per-group runs)."""
    dll = B.load(lib)
    by_shape = {}
    for i, m in enumerate(members):
        by_shape.setdefault(m.shape, []).append(i)
    runs = []
    for shape, idx in by_shape.items():
        run = GroupRun(shape, idx)
        run.group = B.Group(lib, members[idx[0]].prm, len(idx))
        for slot, i in enumerate(idx):
            run.group.init(slot, members[i].prm)
            run.group.set_state(slot, *initial_state(dll, members[i]))
        run.group.graphs(graphs)
        run.group.bootstrap(list(range(len(idx))))
        runs.append(run)

    def work(run):
        try:
            t0 = time.perf_counter()
            run.group.step(list(range(len(run.members))), nsteps)
            run.seconds = time.perf_counter() - t0
        except Exception as e:  # reported by the caller
            run.error = e

    threads = [threading.Thread(target=work, args=(r,)) for r in runs]
    t0 = time.perf_counter()
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    wall = time.perf_counter() - t0
    for r in runs:
        if r.error:
            raise r.error
        if keep_states:
            for slot, i in enumerate(r.members):
                r.states[i] = r.group.get_state(slot)
        r.group.close()
        r.group = None
    return wall, runs


def run_alone(members, nsteps, lib=B.HIP_LIB, keep_states=False, graphs=False):
    """The baseline: each member alone, one after the other (bootstrap untimed).
    graphs=False: plain single simulations (rpcf_step, one launch per kernel);
    graphs=True: each member as a group of one replaying a captured HIP graph per step, so
    that the comparison with run_grouped isolates the batching of members."""
    dll = B.load(lib)
    total = 0.0
    per = []
    states = {}
    for i, m in enumerate(members):
        if graphs:
            with B.Group(lib, m.prm, 1) as g:
                g.init(0, m.prm)
                g.set_state(0, *initial_state(dll, m))
                g.graphs(True)
                g.bootstrap([0])
                t0 = time.perf_counter()
                g.step([0], nsteps)
                dt = time.perf_counter() - t0
                if keep_states:
                    states[i] = g.get_state(0)
        else:
            with B.Simulation(lib, m.prm) as sim:
                sim.set_state(*initial_state(dll, m))
                sim.bootstrap()
                t0 = time.perf_counter()
                sim.step(nsteps)
                dt = time.perf_counter() - t0
                if keep_states:
                    states[i] = sim.get_state()
        total += dt
        per.append(dt)
    return total, per, states


def grid_points(m):
    s = B.sizes(B.load(B.HIP_LIB), m.prm)
    return s["nx"] * s["nz"] * s["ny"]
