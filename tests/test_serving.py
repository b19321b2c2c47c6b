"""Ensembles and the serving loop (rpcf/ensemble.py, rpcf/serving.py) on the group API:
requests arriving while others run are admitted into free slots between chunks of grouped
steps, and every result equals the same simulation run alone (bit for bit).  The request
generators are checked on the CPU."""
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
