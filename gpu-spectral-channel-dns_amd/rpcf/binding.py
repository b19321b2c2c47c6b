"""ctypes binding of the rpcf C ABI (include/rpcf.h, include/rpcf_group.h).

The same binding drives the HIP library (librpcf_hip.so) and the CPU FP64
oracle (librpcf_oracle.so); both export the identical symbol set, so a parity
test runs one script against two libraries.

Spectral arrays are numpy complex128 of shape (P, nz) with P = hnx*ny pencils
in the reference ZXY order p = ky*hnx + kx (reference source/linear.cpp:18).
Physical arrays are float64 of shape (planes, my, mx).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(_HERE)
REPO_ROOT = os.path.dirname(PKG_ROOT)

HIP_LIB = os.path.join(PKG_ROOT, "_build", "librpcf_hip.so")
ORACLE_LIB = os.path.join(REPO_ROOT, "oracle", "_build", "librpcf_oracle.so")

FIELD_U, FIELD_V, FIELD_W, FIELD_OX, FIELD_OY, FIELD_OZ = range(6)


class RpcfError(RuntimeError):
    pass


class Params(ctypes.Structure):
    """Mirror of rpcf_params (include/rpcf.h)."""

    _fields_ = [
        ("mx", ctypes.c_int),
        ("my", ctypes.c_int),
        ("mz", ctypes.c_int),
        ("n_pi_x", ctypes.c_double),
        ("n_pi_y", ctypes.c_double),
        ("Re", ctypes.c_double),
        ("Ro", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("u_profile", ctypes.c_double * 3),
        ("flow_control", ctypes.c_int),
        ("precision", ctypes.c_int),
    ]


def make_params(mx, my, mz, n_pi_x=4.0, n_pi_y=2.0, Re=2600.0, Ro=0.015, dt=0.005,
                u_profile=(0.5, 0.5, 0.0), flow_control=0, precision=0):
    p = Params()
    p.mx, p.my, p.mz = int(mx), int(my), int(mz)
    p.n_pi_x, p.n_pi_y = float(n_pi_x), float(n_pi_y)
    p.Re, p.Ro, p.dt = float(Re), float(Ro), float(dt)
    for i in range(3):
        p.u_profile[i] = float(u_profile[i])
    p.flow_control = int(flow_control)
    p.precision = int(precision)
    return p


_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


def _ptr(a):
    return a.ctypes.data_as(_dp)


_CACHE = {}


def load(path):
    """Load (once) a library exporting the rpcf ABI and declare its signatures."""
    path = os.path.abspath(path)
    if path in _CACHE:
        return _CACHE[path]
    if not os.path.exists(path):
        raise RpcfError(f"library not built: {path}")
    dll = ctypes.CDLL(path)
    P = ctypes.POINTER(Params)
    sig = {
        "rpcf_sizes": (ctypes.c_int, [P] + [ctypes.POINTER(ctypes.c_int)] * 5),
        "rpcf_create": (ctypes.c_int, [P, ctypes.POINTER(_vp)]),
        "rpcf_destroy": (None, [_vp]),
        "rpcf_last_error": (ctypes.c_char_p, [_vp]),
        "rpcf_set_state": (ctypes.c_int, [_vp, _dp, _dp, ctypes.c_long]),
        "rpcf_get_state": (ctypes.c_int, [_vp, _dp, _dp, ctypes.c_long]),
        "rpcf_set_field": (ctypes.c_int, [_vp, ctypes.c_int, _dp, ctypes.c_long]),
        "rpcf_get_field": (ctypes.c_int, [_vp, ctypes.c_int, _dp, ctypes.c_long]),
        "rpcf_get_nonlinear": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, ctypes.c_long]),
        "rpcf_set_nonlinear": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, ctypes.c_long]),
        "rpcf_nonlinear": (ctypes.c_int, [_vp]),
        "rpcf_linear": (ctypes.c_int, [_vp]),
        "rpcf_velocity": (ctypes.c_int, [_vp]),
        "rpcf_step": (ctypes.c_int, [_vp, ctypes.c_int]),
        "rpcf_bootstrap": (ctypes.c_int, [_vp]),
        "rpcf_phys_to_spec": (ctypes.c_int, [_vp, _dp, _dp, ctypes.c_long, ctypes.c_int]),
        "rpcf_spec_to_phys": (ctypes.c_int, [_vp, _dp, _dp, ctypes.c_long, ctypes.c_int]),
        "rpcf_seed_state": (ctypes.c_int, [P, ctypes.c_uint64, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_double, ctypes.c_int, _dp, _dp, ctypes.c_long]),
        "rpcf_get_inverse": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp]),
    }
    ip = ctypes.POINTER(ctypes.c_int)
    group_sig = {  # HIP library only
        "rpcf_group_create": (ctypes.c_int, [P, ctypes.c_int, ctypes.POINTER(_vp)]),
        "rpcf_group_capacity": (ctypes.c_int, [_vp]),
        "rpcf_slot_bytes": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_size_t)]),
        "rpcf_member_init": (ctypes.c_int, [_vp, ctypes.c_int, P]),
        "rpcf_member_release": (ctypes.c_int, [_vp, ctypes.c_int]),
        "rpcf_member_set_state": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, ctypes.c_long]),
        "rpcf_member_get_state": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp, ctypes.c_long]),
        "rpcf_group_bootstrap": (ctypes.c_int, [_vp, ip, ctypes.c_int]),
        "rpcf_group_step": (ctypes.c_int, [_vp, ip, ctypes.c_int, ctypes.c_int]),
        "rpcf_group_graphs": (ctypes.c_int, [_vp, ctypes.c_int]),
        "rpcf_group_cache_hits": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_long)]),
        "rpcf_member_adopt": (ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int]),
        # include/rpcf_static.h: device arrays, no allocation
        "rpcf_spec_to_phys_static": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_long, ctypes.c_int]),
    }
    for name, (res, args) in list(sig.items()) + list(group_sig.items()):
        fn = getattr(dll, name, None)
        if fn is None:
            if name in group_sig:
                continue
            raise RpcfError(f"{path} does not export {name}")
        fn.restype = res
        fn.argtypes = args
    _CACHE[path] = dll
    return dll


def sizes(dll, prm):
    out = [ctypes.c_int() for _ in range(5)]
    rc = dll.rpcf_sizes(ctypes.byref(prm), *[ctypes.byref(o) for o in out])
    if rc != 0:
        raise RpcfError(f"rpcf_sizes: unsupported sizes (rc={rc})")
    nx, ny, nz, pz, hnx = (o.value for o in out)
    return dict(nx=nx, ny=ny, nz=nz, pz=pz, hnx=hnx, P=hnx * ny)


def seed_state(dll, prm, seed, amplitude, kmax, degmax, kexp=0.0, peak_bound=True):
    """rpcf_seed_state: peak_bound=False multiplies the Gaussian coefficients by `amplitude`
    (the stated recipe); True scales the field so that `amplitude` bounds its peak velocity"""
    s = sizes(dll, prm)
    v = np.zeros((s["P"], s["nz"]), np.complex128)
    e = np.zeros((s["P"], s["nz"]), np.complex128)
    rc = dll.rpcf_seed_state(ctypes.byref(prm), seed, amplitude, kmax, degmax, kexp, int(bool(peak_bound)),
                             _ptr(v.view(np.float64)), _ptr(e.view(np.float64)), s["nz"])
    if rc != 0:
        raise RpcfError(f"rpcf_seed_state rc={rc}")
    return v, e


class Simulation:
    """One rpcf context (one simulation) behind either library."""

    def __init__(self, lib_path, prm):
        self.dll = load(lib_path)
        self.prm = prm
        self.s = sizes(self.dll, prm)
        h = _vp()
        rc = self.dll.rpcf_create(ctypes.byref(prm), ctypes.byref(h))
        if rc != 0:
            raise RpcfError(f"rpcf_create failed rc={rc}")
        self.h = h

    # -- helpers ------------------------------------------------------------
    def _check(self, rc, what):
        if rc != 0:
            msg = self.dll.rpcf_last_error(self.h)
            raise RpcfError(f"{what}: rc={rc}: {msg.decode() if msg else ''}")

    def _spec(self):
        return np.zeros((self.s["P"], self.s["nz"]), np.complex128)

    def close(self):
        if getattr(self, "h", None):
            self.dll.rpcf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state --------------------------------------------------------------
    def set_state(self, v, eta):
        v = np.ascontiguousarray(v, np.complex128)
        eta = np.ascontiguousarray(eta, np.complex128)
        self._check(self.dll.rpcf_set_state(self.h, _ptr(v.view(np.float64)), _ptr(eta.view(np.float64)),
                                            v.shape[1]), "set_state")

    def get_state(self):
        v, e = self._spec(), self._spec()
        self._check(self.dll.rpcf_get_state(self.h, _ptr(v.view(np.float64)), _ptr(e.view(np.float64)),
                                            self.s["nz"]), "get_state")
        return v, e

    def set_field(self, which, a):
        a = np.ascontiguousarray(a, np.complex128)
        self._check(self.dll.rpcf_set_field(self.h, which, _ptr(a.view(np.float64)), a.shape[1]), "set_field")

    def get_field(self, which):
        a = self._spec()
        self._check(self.dll.rpcf_get_field(self.h, which, _ptr(a.view(np.float64)), self.s["nz"]), "get_field")
        return a

    def get_nonlinear(self, prev=False):
        v, e = self._spec(), self._spec()
        self._check(self.dll.rpcf_get_nonlinear(self.h, int(prev), _ptr(v.view(np.float64)),
                                                _ptr(e.view(np.float64)), self.s["nz"]), "get_nonlinear")
        return v, e

    def set_nonlinear(self, n_v, n_eta, prev=False):
        n_v = np.ascontiguousarray(n_v, np.complex128)
        n_eta = np.ascontiguousarray(n_eta, np.complex128)
        self._check(self.dll.rpcf_set_nonlinear(self.h, int(prev), _ptr(n_v.view(np.float64)),
                                                _ptr(n_eta.view(np.float64)), n_v.shape[1]), "set_nonlinear")

    # -- stepping -----------------------------------------------------------
    def nonlinear(self):
        self._check(self.dll.rpcf_nonlinear(self.h), "nonlinear")

    def linear(self):
        self._check(self.dll.rpcf_linear(self.h), "linear")

    def velocity(self):
        self._check(self.dll.rpcf_velocity(self.h), "velocity")

    def step(self, n=1):
        self._check(self.dll.rpcf_step(self.h, int(n)), "step")

    def bootstrap(self):
        self._check(self.dll.rpcf_bootstrap(self.h), "bootstrap")

    # -- transforms ---------------------------------------------------------
    def phys_to_spec(self, phys, padded=True):
        phys = np.ascontiguousarray(phys, np.float64)
        out = self._spec()
        self._check(self.dll.rpcf_phys_to_spec(self.h, _ptr(phys), _ptr(out.view(np.float64)), self.s["nz"],
                                               int(padded)), "phys_to_spec")
        return out

    def spec_to_phys(self, spec, padded=True):
        spec = np.ascontiguousarray(spec, np.complex128)
        planes = self.s["pz"] if padded else self.s["nz"]
        out = np.zeros((planes, self.prm.my, self.prm.mx), np.float64)
        self._check(self.dll.rpcf_spec_to_phys(self.h, _ptr(spec.view(np.float64)), _ptr(out), spec.shape[1],
                                               int(padded)), "spec_to_phys")
        return out

    def get_inverse(self, eq, pencil):
        nz = self.s["nz"]
        out = np.zeros((nz, nz), np.complex128)
        self._check(self.dll.rpcf_get_inverse(self.h, int(eq), int(pencil), _ptr(out.view(np.float64))),
                    "get_inverse")
        return out


def slot_bytes(dll, prm):
    """device bytes one group slot of this grid shape occupies"""
    b = ctypes.c_size_t()
    if dll.rpcf_slot_bytes(ctypes.byref(prm), ctypes.byref(b)) != 0:
        raise RpcfError("rpcf_slot_bytes: unsupported sizes")
    return b.value


class Group:
    """A group of simulations of one grid shape advanced by grouped launches
    (include/rpcf_group.h, HIP library only)."""

    def __init__(self, lib_path, shape, capacity):
        self.dll = load(lib_path)
        self.shape = shape
        self.s = sizes(self.dll, shape)
        h = _vp()
        rc = self.dll.rpcf_group_create(ctypes.byref(shape), int(capacity), ctypes.byref(h))
        if rc != 0:
            raise RpcfError(f"rpcf_group_create failed rc={rc}")
        self.h = h
        self.capacity = int(capacity)

    def _check(self, rc, what):
        if rc != 0:
            msg = self.dll.rpcf_last_error(self.h)
            raise RpcfError(f"{what}: rc={rc}: {msg.decode() if msg else ''}")

    def _spec(self):
        return np.zeros((self.s["P"], self.s["nz"]), np.complex128)

    def close(self):
        if getattr(self, "h", None):
            self.dll.rpcf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def init(self, slot, prm):
        self._check(self.dll.rpcf_member_init(self.h, int(slot), ctypes.byref(prm)), "member_init")

    def release(self, slot):
        self._check(self.dll.rpcf_member_release(self.h, int(slot)), "member_release")

    def set_state(self, slot, v, eta):
        v = np.ascontiguousarray(v, np.complex128)
        eta = np.ascontiguousarray(eta, np.complex128)
        self._check(self.dll.rpcf_member_set_state(self.h, int(slot), _ptr(v.view(np.float64)),
                                                   _ptr(eta.view(np.float64)), v.shape[1]), "member_set_state")

    def get_state(self, slot):
        v, e = self._spec(), self._spec()
        self._check(self.dll.rpcf_member_get_state(self.h, int(slot), _ptr(v.view(np.float64)),
                                                   _ptr(e.view(np.float64)), self.s["nz"]), "member_get_state")
        return v, e

    @staticmethod
    def _ids(slots):
        a = (ctypes.c_int * len(slots))(*[int(x) for x in slots])
        return a, len(slots)

    def bootstrap(self, slots):
        a, n = self._ids(slots)
        self._check(self.dll.rpcf_group_bootstrap(self.h, a, n), "group_bootstrap")

    def step(self, slots, nsteps=1):
        a, n = self._ids(slots)
        self._check(self.dll.rpcf_group_step(self.h, a, n, int(nsteps)), "group_step")

    def graphs(self, enable=True):
        self._check(self.dll.rpcf_group_graphs(self.h, int(bool(enable))), "group_graphs")

    def adopt(self, slot, src, src_slot=0):
        """move the member of `src` (another group of this shape and precision) into `slot`:
        device-to-device copies of its operators, inverses, state and history"""
        self._check(self.dll.rpcf_member_adopt(self.h, int(slot), src.h, int(src_slot)), "member_adopt")

    def cache_hits(self):
        """member initialisations whose inverses were copied from another slot"""
        n = ctypes.c_long()
        self._check(self.dll.rpcf_group_cache_hits(self.h, ctypes.byref(n)), "group_cache_hits")
        return n.value
