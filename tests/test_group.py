"""Groups (include/rpcf_group.h): several simulations of one grid shape advanced by
grouped launches.  A member's result in the mix is bit-identical to the same simulation
run alone (each kernel computes a member's values with the same arithmetic whatever else
is in the launch), with and without HIP-graph replay; subsets, release and re-use of slots
leave the other members untouched.  GPU only (the oracle runs one simulation per
context); one member is also checked against the oracle."""
import numpy as np
import pytest

import rpcf
from helpers import rel

pytestmark = pytest.mark.gpu

SHAPE = dict(mx=48, my=48, mz=128)
MEMBERS = [
    dict(n_pi_x=2.0, n_pi_y=1.0, Re=1000.0, Ro=0.0, dt=0.005, u_profile=(0.0, 1.0, 0.0), flow_control=0),
    dict(n_pi_x=4.0, n_pi_y=2.0, Re=2600.0, Ro=0.3, dt=0.004, u_profile=(1.0, 0.0, -1.0), flow_control=1),
    dict(n_pi_x=2.0, n_pi_y=4.0 / 3.0, Re=700.0, Ro=0.1, dt=0.002, u_profile=(0.5, 0.5, 0.0), flow_control=0),
]
IC = (0.02, 4, 8, 0.0)


def member_params(i):
    return rpcf.make_params(**SHAPE, **MEMBERS[i])


def initial(i):
    prm = member_params(i)
    return rpcf.seed_state(rpcf.load(rpcf.HIP_LIB), prm, 100 + i, *IC)


def alone(i, nsteps):
    with rpcf.Simulation(rpcf.HIP_LIB, member_params(i)) as sim:
        sim.set_state(*initial(i))
        sim.bootstrap()
        sim.step(nsteps)
        return sim.get_state()


def make_group(graphs=False):
    g = rpcf.Group(rpcf.HIP_LIB, rpcf.make_params(**SHAPE), len(MEMBERS))
    for i in range(len(MEMBERS)):
        g.init(i, member_params(i))
        g.set_state(i, *initial(i))
    g.graphs(graphs)
    return g


@pytest.mark.parametrize("graphs", [False, True])
def test_members_match_alone(graphs):
    nsteps = 6
    with make_group(graphs) as g:
        slots = list(range(len(MEMBERS)))
        g.bootstrap(slots)
        for _ in range(3):  # several replays with host work in between
            g.step(slots, 2)
            _ = g.get_state(0)
        got = [g.get_state(i) for i in slots]
    for i in range(len(MEMBERS)):
        ref = alone(i, nsteps)
        assert np.array_equal(got[i][0], ref[0]) and np.array_equal(got[i][1], ref[1]), i


def test_member_matches_oracle():
    nsteps = 3
    with make_group() as g:
        g.bootstrap([0, 1, 2])
        g.step([0, 1, 2], nsteps)
        v, e = g.get_state(1)
    with rpcf.Simulation(rpcf.ORACLE_LIB, member_params(1)) as o:
        o.set_state(*initial(1))
        o.bootstrap()
        o.step(nsteps)
        vo, eo = o.get_state()
    assert rel(v, vo) < 1e-11 and rel(e, eo) < 1e-11


def test_subset_release_and_reuse():
    with make_group(graphs=True) as g:
        g.bootstrap([0, 1, 2])
        s1 = g.get_state(1)
        g.step([0, 2], 3)  # slot 1 stays where it was
        v1, e1 = g.get_state(1)
        assert np.array_equal(v1, s1[0]) and np.array_equal(e1, s1[1])
        g.release(1)
        with pytest.raises(rpcf.RpcfError):
            g.step([0, 1], 1)
        # slot 1 takes member 2's physics and initial field: it then tracks member 2 run alone
        g.init(1, member_params(2))
        g.set_state(1, *initial(2))
        g.bootstrap([1])
        g.step([1, 0], 3)
        v, e = g.get_state(1)
    ref = alone(2, 3)
    assert np.array_equal(v, ref[0]) and np.array_equal(e, ref[1])


def test_group_errors():
    with rpcf.Group(rpcf.HIP_LIB, rpcf.make_params(**SHAPE), 2) as g:
        with pytest.raises(rpcf.RpcfError):
            g.init(2, member_params(0))  # outside capacity
        with pytest.raises(rpcf.RpcfError):
            g.init(0, rpcf.make_params(mx=96, my=48, mz=128))  # another shape
        g.init(0, member_params(0))
        with pytest.raises(rpcf.RpcfError):
            g.step([0], 1)  # no state
        g.set_state(0, *initial(0))
        with pytest.raises(rpcf.RpcfError):
            g.step([0, 0], 1)  # listed twice
        with pytest.raises(rpcf.RpcfError):
            g.step([1], 1)  # empty slot


def test_rotation_sweep_shares_inverses():
    """members that differ only in Ro, flow control and seed take their inverses from the
    matrix cache (copied from the first slot) and still match the same simulations alone"""
    base = dict(MEMBERS[0])
    sweep = [dict(base, Ro=ro, flow_control=fc) for ro, fc in ((0.0, 0), (0.15, 0), (0.3, 1))]
    prms = [rpcf.make_params(**SHAPE, **m) for m in sweep]
    states = [rpcf.seed_state(rpcf.load(rpcf.HIP_LIB), p, 200 + i, *IC) for i, p in enumerate(prms)]
    with rpcf.Group(rpcf.HIP_LIB, prms[0], len(prms)) as g:
        for i, p in enumerate(prms):
            g.init(i, p)
            g.set_state(i, *states[i])
        assert g.cache_hits() == len(prms) - 1
        slots = list(range(len(prms)))
        g.bootstrap(slots)
        g.step(slots, 4)
        got = [g.get_state(i) for i in slots]
    for i, p in enumerate(prms):
        with rpcf.Simulation(rpcf.HIP_LIB, p) as sim:
            sim.set_state(*states[i])
            sim.bootstrap()
            sim.step(4)
            ref = sim.get_state()
        assert np.array_equal(got[i][0], ref[0]) and np.array_equal(got[i][1], ref[1]), i
