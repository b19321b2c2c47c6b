"""rpcf_spec_to_phys_static (include/rpcf_static.h): the backward transform of one field
with device arrays and no allocation, its intermediates in the library's static device
scratch.  Checked against rpcf_spec_to_phys (which allocates its intermediates per call)
on the same inputs: the same kernels run, so the values must be bit-identical.  Two calls
in a row, the second on a smaller grid (the scratch keeps the first call's contents), a
pencil stride larger than nz, and the rejection of a transform whose intermediates exceed
the scratch.  GPU tests."""
import ctypes

import numpy as np
import pytest

import rpcf
from helpers import IC, params

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


def _hip():
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


class DeviceArray:
    """device memory of the test (the entry point under test allocates nothing itself)"""

    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), nbytes) == 0
        self.nbytes = nbytes

    def upload(self, a):
        assert a.nbytes == self.nbytes
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, _H2D) == 0

    def download(self, a):
        assert a.nbytes == self.nbytes
        assert self.hip.hipMemcpy(a.ctypes.data, self.p, a.nbytes, _D2H) == 0

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.hip.hipFree(self.p)


def static_transform(sim, spec, padded):
    """spec: (P, stride) complex128 with the coefficients in [:, :nz]"""
    hip = _hip()
    planes = sim.s["pz"] if padded else sim.s["nz"]
    out = np.full((planes, sim.prm.my, sim.prm.mx), np.nan)
    with DeviceArray(hip, spec.nbytes) as ds, DeviceArray(hip, out.nbytes) as dp:
        ds.upload(np.ascontiguousarray(spec))
        dp.upload(out)
        rc = sim.dll.rpcf_spec_to_phys_static(sim.h, ds.p, dp.p, spec.shape[1], int(padded))
        sim._check(rc, "spec_to_phys_static")
        dp.download(out)
    return out


def seeded_field(sim, name, seed):
    v, _ = rpcf.seed_state(sim.dll, sim.prm, seed, *IC[name])
    return v


@pytest.mark.parametrize("padded", [1, 0])
def test_static_matches_allocating_entry_point(padded):
    """two calls in a row on one GPU, the second on a smaller grid: each bit-identical to
    rpcf_spec_to_phys on the same field"""
    for name, seed in (("REFCOUETTE", 3), ("TINY", 4)):
        with rpcf.Simulation(rpcf.HIP_LIB, params(name)) as sim:
            spec = seeded_field(sim, name, seed)
            want = sim.spec_to_phys(spec, padded=bool(padded))
            got = static_transform(sim, spec, padded)
            print(name, "padded" if padded else "unpadded", "max |static - allocating| =",
                  float(np.abs(got - want).max()), "max |field| =", float(np.abs(want).max()))
            assert np.abs(want).max() > 0
            assert np.array_equal(got, want)


def test_static_pencil_stride():
    """a spectral array whose pencils are further apart than nz"""
    with rpcf.Simulation(rpcf.HIP_LIB, params("TINY")) as sim:
        spec = seeded_field(sim, "TINY", 7)
        nz = sim.s["nz"]
        wide = np.full((spec.shape[0], nz + 3), 1e300 + 0j)
        wide[:, :nz] = spec
        want = sim.spec_to_phys(spec, padded=True)
        assert np.array_equal(static_transform(sim, wide, 1), want)


def test_static_rejects_sizes_beyond_the_scratch():
    """128x129x128 (mx = my = 192, mz = 512): unpadded fits the 64 MiB scratch, padded
    does not and is refused with a message naming both sizes"""
    prm = params("C1")
    with rpcf.Simulation(rpcf.HIP_LIB, prm) as sim:
        spec = np.zeros((sim.s["P"], sim.s["nz"]), np.complex128)
        spec[1, 0] = 1.0
        with pytest.raises(rpcf.RpcfError, match="bytes of scratch, the maximum is 67108864"):
            static_transform(sim, spec, 1)
        got = static_transform(sim, spec, 0)
        assert np.array_equal(got, sim.spec_to_phys(spec, padded=False))
