"""The benchmark configurations as stated, against the CPU oracle (GPU tests).

  C0  32x33x32 plane Couette, U = y, Re = 1000, dt = 0.005, the stated initial field
      (Gaussian coefficients x 1e-2, |kx|, |kz| <= 4, degree <= 8, seed 0): FP64 within
      1e-12 of the oracle after one step and 1e-9 after 100; the mixed form within 1e-6 of
      the oracle after 100 steps on every field.
  C1  128x129x128 channel, constant flow rate: the same bars after 1 and 100 steps.  Its
      literal initial field (coefficients x 0.1, |k| <= 16, degree <= 24, x (1+|k|)^(-5/6))
      blows up within the first steps at dt = 0.002 on either library (shown below), so the
      field is scaled to a peak perturbation velocity of 0.1.
  C2  256x97x256 rotating plane Couette (Ro = 0.1, Re = 1300): the same bars after 1 and 20
      steps (the oracle takes about half a second per step on this grid); its field is scaled
      to a peak perturbation velocity of 0.05 for the same reason as C1's.

Per field: v-hat and eta-hat of the state, and the six velocity / vorticity fields recovered
from it; metric max|a - b| / max|b| over all pencils.
"""
import numpy as np
import pytest

import rpcf
from helpers import IC, params, rel

pytestmark = pytest.mark.gpu


def fields_of(sim):
    v, e = sim.get_state()
    return [v, e] + [sim.get_field(f) for f in range(6)]


NAMES = ["v", "eta", "u", "w_span", "v_wall", "omega_x", "omega_span", "omega_wall"]


def compare(a, b):
    return {n: rel(x, y) for n, x, y in zip(NAMES, a, b) if np.abs(y).max() > 0}


def run_against_oracle(name, steps):
    prm = params(name)
    ic = IC[name]
    ref = rpcf.Simulation(rpcf.ORACLE_LIB, prm)
    v0, e0 = rpcf.seed_state(ref.dll, prm, 0, *ic)
    out = {}
    sims = {"oracle": ref, "fp64": rpcf.Simulation(rpcf.HIP_LIB, prm)}
    mp = params(name, precision=1)
    sims["mixed"] = rpcf.Simulation(rpcf.HIP_LIB, mp)
    for key, sim in sims.items():
        sim.set_state(v0, e0)
        sim.bootstrap()
        out[(key, 1)] = fields_of(sim)
        sim.step(steps - 1)
        out[(key, steps)] = fields_of(sim)
        sim.close()
    return out


@pytest.mark.parametrize("name", ["C0", "C1", "C2"])
def test_configuration_against_oracle(name):
    steps = 20 if name == "C2" else 100
    out = run_against_oracle(name, steps)
    e1 = compare(out[("fp64", 1)], out[("oracle", 1)])
    e100 = compare(out[("fp64", steps)], out[("oracle", steps)])
    em = compare(out[("mixed", steps)], out[("oracle", steps)])
    print(name, "FP64 vs oracle after 1 step:", {k: "%.1e" % v for k, v in e1.items()})
    print(name, "FP64 vs oracle after %d steps:" % steps, {k: "%.1e" % v for k, v in e100.items()})
    print(name, "mixed vs oracle after %d steps:" % steps, {k: "%.1e" % v for k, v in em.items()})
    assert max(e1.values()) < 1e-12
    assert max(e100.values()) < 1e-9
    assert max(em.values()) < 1e-6
    assert max(em.values()) > 1e-13  # the FP32 path ran


def test_c1_literal_initial_field_blows_up():
    """the channel configuration's initial field read literally (coefficients x 0.1) has a
    peak velocity of order 30 against the unit centreline velocity, CFL about 1.2 at
    dt = 0.002: the FP64 GPU step blows up within 20 steps, which is why the C1 workload
    scales its field to a peak perturbation velocity of 0.1"""
    prm = params("C1")
    with rpcf.Simulation(rpcf.HIP_LIB, prm) as sim:
        amp, kmax, deg, kexp, _ = IC["C1"]
        v0, e0 = rpcf.seed_state(sim.dll, prm, 0, amp, kmax, deg, kexp, peak_bound=False)
        sim.set_state(v0, e0)
        sim.bootstrap()
        blown = None
        for n in range(2, 21):
            sim.step(1)
            v, e = sim.get_state()
            if not (np.isfinite(v).all() and np.isfinite(e).all()) or np.abs(v[1:]).max() > 1e3 * np.abs(v0).max():
                blown = n
                break
    print("C1 literal field: blow-up at step", blown)
    assert blown is not None
