"""The C ABI boundary: both libraries load and export every entry point declared in
include/rpcf.h; argument checking and error reporting work without a GPU (calls that
need no device only).  CPU only."""
import ctypes
import os
import re

import numpy as np
import pytest

import rpcf
from helpers import IC, params

INCLUDE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
HEADER = os.path.join(INCLUDE, "rpcf.h")
GROUP_HEADER = os.path.join(INCLUDE, "rpcf_group.h")
STATIC_HEADER = os.path.join(INCLUDE, "rpcf_static.h")


def declared(header=HEADER):
    src = open(header).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(rpcf_\w+)\s*\(", src)))


def test_header_declares_the_step_api():
    names = declared()
    for n in ["rpcf_create", "rpcf_destroy", "rpcf_set_state", "rpcf_get_state", "rpcf_nonlinear", "rpcf_linear",
              "rpcf_velocity", "rpcf_step", "rpcf_bootstrap", "rpcf_last_error", "rpcf_seed_state"]:
        assert n in names


@pytest.mark.parametrize("which", ["oracle", "hip"])
def test_library_exports_every_declared_symbol(which):
    path = rpcf.ORACLE_LIB if which == "oracle" else rpcf.HIP_LIB
    dll = ctypes.CDLL(path)
    missing = [n for n in declared() if not hasattr(dll, n)]
    assert not missing, missing


def test_hip_library_exports_the_group_api():
    """include/rpcf_group.h is implemented by the HIP library (the oracle runs one
    simulation per context)"""
    dll = ctypes.CDLL(rpcf.HIP_LIB)
    names = declared(GROUP_HEADER)
    assert "rpcf_group_step" in names and "rpcf_member_init" in names
    missing = [n for n in names if not hasattr(dll, n)]
    assert not missing, missing


def test_hip_library_exports_the_static_api():
    """include/rpcf_static.h (entry points that allocate nothing) is implemented by the HIP
    library"""
    dll = ctypes.CDLL(rpcf.HIP_LIB)
    names = declared(STATIC_HEADER)
    assert names == ["rpcf_spec_to_phys_static"]
    missing = [n for n in names if not hasattr(dll, n)]
    assert not missing, missing


def test_slot_bytes():
    """device bytes of one slot: the per-slot arrays, which the inverses dominate -- the full
    pair, their parity blocks (nz x (N/2 + 1) complex FP64 each) and, in the FP64 form, the
    BF16 remainders (4 B per entry) of the v blocks"""
    dll = rpcf.load(rpcf.HIP_LIB)
    prm = params("C1")
    b = rpcf.slot_bytes(dll, prm)
    s = rpcf.sizes(dll, prm)
    nslot = s["hnx"] * (s["ny"] // 2 + 1)
    np_ = (s["nz"] - 1) // 2 + 1
    minv = 2 * nslot * s["nz"] ** 2 * 16
    blocks = 2 * nslot * s["nz"] * np_ * 16
    rem = nslot * s["nz"] * np_ * 4
    assert minv + blocks + rem < b < 1.35 * (minv + blocks + rem)  # the fields add ~1 GB
    prm.precision = 2  # the parity blocks alone
    b2 = rpcf.slot_bytes(dll, prm)
    assert b - b2 == rem
    prm.precision = 1  # the eta blocks as FP32 high and BF16 low parts (12 B per entry) only
    b1 = rpcf.slot_bytes(dll, prm)
    assert b1 - b2 == 3 * rem - blocks // 2
    with pytest.raises(rpcf.RpcfError):
        rpcf.slot_bytes(dll, params("C0", mz=130))


@pytest.mark.parametrize("which", ["oracle", "hip"])
def test_sizes_and_rejections(which):
    dll = rpcf.load(rpcf.ORACLE_LIB if which == "oracle" else rpcf.HIP_LIB)
    s = rpcf.sizes(dll, params("C1"))
    assert (s["nx"], s["ny"], s["nz"], s["pz"], s["hnx"]) == (128, 128, 129, 257, 65)
    s = rpcf.sizes(dll, params("REF128"))
    assert (s["nx"], s["ny"], s["nz"], s["pz"]) == (84, 84, 33, 65)
    for bad in (dict(mz=130), dict(mx=47), dict(Re=0.0), dict(dt=-1.0)):
        with pytest.raises(rpcf.RpcfError):
            rpcf.sizes(dll, params("C0", **bad))
    out = ctypes.c_void_p()
    assert dll.rpcf_create(ctypes.byref(params("C0", mz=130)), ctypes.byref(out)) == -1
    assert dll.rpcf_create(None, ctypes.byref(out)) == -1


def test_seed_identical_in_both_libraries():
    """the HIP library's host-side generator and the oracle's restatement agree bit for bit"""
    for name in ("C0", "C1", "TINY"):
        prm = params(name)
        a = rpcf.seed_state(rpcf.load(rpcf.ORACLE_LIB), prm, 5, *IC[name])
        b = rpcf.seed_state(rpcf.load(rpcf.HIP_LIB), prm, 5, *IC[name])
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_seed_rejects_too_high_degree():
    prm = params("TINY")
    for path in (rpcf.ORACLE_LIB, rpcf.HIP_LIB):
        with pytest.raises(rpcf.RpcfError):
            rpcf.seed_state(rpcf.load(path), prm, 0, 1.0, 4, 13, 0.0)  # nz = 17: degree + 4 must be < nz


def test_oracle_errors_are_reported_not_fatal():
    prm = params("TINY")
    with rpcf.Simulation(rpcf.ORACLE_LIB, prm) as sim:
        with pytest.raises(rpcf.RpcfError, match="no state"):
            sim.step(1)
        with pytest.raises(rpcf.RpcfError, match="bad field"):
            sim.get_field(7)
        z = np.zeros((sim.s["P"], sim.s["nz"]), np.complex128)
        sim.set_state(z, z)
        sim.step(1)
