"""Ensembles and the serving loop (rpcf/ensemble.py, rpcf/serving.py) on the group API:
requests arriving while others run are staged off the running groups' streams and adopted
into free slots between chunks of grouped steps, and every result equals the same
simulation run alone (bit for bit).  The request generators are checked on the CPU."""
import numpy as np
import pytest

import rpcf
from rpcf import ensemble, serving


def test_mixed_members_recipe():
    ms = ensemble.mixed_members(64, seed=0)
    assert len(ms) == 64 and [m.seed for m in ms] == list(range(64))
    shapes = {m.shape for m in ms}
    assert shapes <= {(24, 24, 128), (48, 48, 192), (48, 48, 256), (96, 96, 256)} and len(shapes) == 4
    for m in ms:
        assert 500 <= m.prm.Re <= 3000 and 0 <= m.prm.Ro <= 0.5 and m.prm.dt == 0.002
        couette = tuple(m.prm.u_profile) == (0.0, 1.0, 0.0)
        assert couette or (tuple(m.prm.u_profile) == (1.0, 0.0, -1.0) and m.prm.flow_control == 1)
    s = rpcf.sizes(rpcf.load(rpcf.ORACLE_LIB), ms[0].prm)
    assert s["nz"] in (33, 49, 65)


def test_poisson_requests_recipe():
    reqs = serving.poisson_requests(rate=20.0, duration=60.0, seed=1)
    t = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(t) > 0) and t[-1] < 60.0
    assert abs(len(reqs) - 1200) < 5 * np.sqrt(1200)  # Poisson count
    assert all(100 <= r.steps <= 2000 for r in reqs)
    assert {r.shape for r in reqs} == {(48, 48, 128), (96, 96, 256), (192, 192, 512), (384, 384, 512)}


def alone(prm, seed, ic, steps):
    with rpcf.Simulation(rpcf.HIP_LIB, prm) as sim:
        sim.set_state(*rpcf.seed_state(sim.dll, prm, seed, *ic))
        sim.bootstrap()
        sim.step(steps - 1)
        return sim.get_state()


@pytest.mark.gpu
def test_ensemble_grouped_equals_alone():
    ms = ensemble.mixed_members(10, seed=3)
    _, runs = ensemble.run_grouped(ms, 5, keep_states=True)
    got = {i: st for r in runs for i, st in r.states.items()}
    _, _, ref = ensemble.run_alone(ms, 5, keep_states=True)
    for i in range(len(ms)):
        assert np.array_equal(got[i][0], ref[i][0]) and np.array_equal(got[i][1], ref[i][1]), i


@pytest.mark.gpu
def test_serving_results_equal_alone():
    grids = [(32, 33, 32), (64, 65, 64)]
    reqs = serving.poisson_requests(rate=200.0, duration=0.1, seed=7, grids=grids, steps=(8, 40))
    for r in reqs:
        r.keep_state = True
    srv = serving.Server(chunk=7, capacity=lambda shape, nbytes: 3)
    done = srv.run(reqs, horizon=120.0)
    assert len(done) == len(reqs) and all(r.error is None for r in done)
    for r in done:
        assert r.start >= r.arrival and r.end >= r.start and r.done == r.steps
        v, e = alone(r.prm, r.seed, r.ic, r.steps)
        assert np.array_equal(r.state[0], v) and np.array_equal(r.state[1], e), r.rid
    stats = serving.summarize(done, reqs, max(r.end for r in done))
    assert stats["finished"] == len(reqs) and stats["grid_point_steps_per_s"] > 0
    assert stats["unfinished_running"] == 0 and stats["unfinished_queued"] == 0


@pytest.mark.gpu
def test_serving_request_grids_of_the_workload():
    """the serving workload's two large grids, 128x129x128 and 256x129x256 (which the
    reference cannot run: nz > MAX_NZ = 100), admitted into running groups while small
    requests run: every result bit-identical to the simulation alone; a 128x129x128 request
    of one step (its bootstrap) within 1e-12 of the CPU oracle"""
    small = serving.poisson_requests(rate=400.0, duration=0.05, seed=11, grids=[(32, 33, 32)], steps=(30, 60))
    large = serving.poisson_requests(rate=200.0, duration=0.05, seed=12, grids=[(128, 129, 128), (256, 129, 256)],
                                     steps=(2, 4))[:4]
    t_small = max(r.arrival for r in small)
    reqs = []
    for r in small:
        reqs.append(r)
    for k, r in enumerate(large):  # the large ones arrive while the small ones run
        r.arrival = t_small + 0.02 + 0.01 * k
        reqs.append(r)
    shapes = {r.shape for r in large}
    assert shapes == {(192, 192, 512), (384, 384, 512)}, shapes
    one = next(r for r in large if r.shape == (192, 192, 512))
    one.steps = 1
    for i, r in enumerate(reqs):
        r.rid, r.keep_state = i, True
    srv = serving.Server(chunk=5, capacity=lambda shape, nbytes: 2)
    done = srv.run(reqs, horizon=300.0)
    assert len(done) == len(reqs) and all(r.error is None for r in done)
    for r in done:
        assert r.adopted is not None and r.start <= r.adopted <= r.end and r.done == r.steps
        v, e = alone(r.prm, r.seed, r.ic, r.steps)
        assert np.array_equal(r.state[0], v) and np.array_equal(r.state[1], e), (r.rid, r.shape)
    with rpcf.Simulation(rpcf.ORACLE_LIB, one.prm) as o:
        o.set_state(*rpcf.seed_state(o.dll, one.prm, one.seed, *one.ic))
        o.bootstrap()
        ov, oe = o.get_state()
    err = max(float(np.abs(one.state[0] - ov).max() / np.abs(ov).max()),
              float(np.abs(one.state[1] - oe).max() / np.abs(oe).max()))
    print("128x129x128 request after its bootstrap step vs oracle:", err)
    assert err < 1e-12


def test_summary_reports_unfinished_requests():
    """requests still running or queued at the horizon are counted and enter the lower-bound
    percentiles instead of being dropped (no GPU: sizes only)"""
    reqs = serving.poisson_requests(rate=50.0, duration=1.0, seed=2, grids=[(32, 33, 32)], steps=(100, 100))
    done = []
    for k, r in enumerate(reqs):
        if k % 3 == 0:  # finished
            r.start, r.adopted, r.end, r.done = r.arrival + 0.1, r.arrival + 0.2, r.arrival + 1.0, r.steps
            done.append(r)
        elif k % 3 == 1:  # running at the horizon
            r.start, r.adopted, r.done = r.arrival + 0.1, r.arrival + 0.2, 50
    wall = 5.0
    st = serving.summarize(done, reqs, wall)
    assert st["finished"] == len(done)
    assert st["unfinished_running"] == sum(1 for k in range(len(reqs)) if k % 3 == 1)
    assert st["unfinished_queued"] == sum(1 for k in range(len(reqs)) if k % 3 == 2)
    assert st["p99_s_per_100_steps_with_unfinished_lower_bound"] > st["p99_s_per_100_steps"]
    assert st["p99_queue_s_with_unfinished_lower_bound"] > st["p99_queue_s"]
