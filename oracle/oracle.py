"""ctypes front end of the CPU oracle (oracle/liblbm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as what the tests compare against or as the
timed CPU baseline.

``OracleDomain`` holds one (sub)domain in host memory (numpy) and steps it
with the C restatement of the reference kernels; ``OracleMulti`` runs a
decomposition with the reference's exchange semantics
(src/CController.hpp:265-391): after alpha each receiver's ghost layer gets all
19 slots of the neighbour's inner layer, after beta each receiver's inner
layer gets the neighbour's ghost layer for populations with n.e > 0, faces in
x, then y, then z order.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liblbm_oracle.so")

FLAG_OBSTACLE, FLAG_FLUID, FLAG_INJECTION, FLAG_GHOST = 1, 2, 4, 8

E = np.array(
    [[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0],
     [1, 1, 0], [-1, -1, 0], [1, -1, 0], [-1, 1, 0],
     [1, 0, 1], [-1, 0, -1], [1, 0, -1], [-1, 0, 1],
     [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1],
     [0, 0, 1], [0, 0, -1], [0, 0, 0]], dtype=np.int64)
OPP = np.array([i ^ 1 if i < 18 else 18 for i in range(19)])

_libs = {}
_lock = threading.Lock()


class _ParamsF(C.Structure):
    _fields_ = [("inv_tau", C.c_float), ("tau0", C.c_float), ("les_coeff", C.c_float),
                ("gx", C.c_float), ("gy", C.c_float), ("gz", C.c_float), ("lid", C.c_float),
                ("forcing", C.c_int), ("store_velocity", C.c_int), ("store_density", C.c_int)]


class _ParamsD(C.Structure):
    _fields_ = [("inv_tau", C.c_double), ("tau0", C.c_double), ("les_coeff", C.c_double),
                ("gx", C.c_double), ("gy", C.c_double), ("gz", C.c_double), ("lid", C.c_double),
                ("forcing", C.c_int), ("store_velocity", C.c_int), ("store_density", C.c_int)]


def load(path=None):
    """The oracle library: the parity build by default, or another build of
    the same source (the timing build of the CPU baseline) at `path`."""
    path = path or LIB
    with _lock:
        if path in _libs:
            return _libs[path]
        if not os.path.exists(path):
            raise ImportError(f"{path} not found: run python oracle/build_oracle.py")
        lib = C.CDLL(path)
        P = C.c_void_p
        I = C.c_int
        for sfx, params in (("_f", _ParamsF), ("_d", _ParamsD)):
            getattr(lib, "oracle_alpha" + sfx).argtypes = [P, P, P, P, I, I, I, C.POINTER(params)]
            getattr(lib, "oracle_beta" + sfx).argtypes = [P, P, P, P, I, I, I, C.POINTER(params)]
            getattr(lib, "oracle_init" + sfx).argtypes = [P, P, P, P, I, I, I, P, C.c_double, C.c_uint64,
                                                         C.c_uint64, P, P]
            getattr(lib, "oracle_store_dd_region" + sfx).argtypes = [P, P, P, P, P]
            getattr(lib, "oracle_set_dd_region" + sfx).argtypes = [P, P, P, P, P, P]
            getattr(lib, "oracle_copy_dd_box" + sfx).argtypes = [P, P, P, P, P, P, P, P]
            getattr(lib, "oracle_moments" + sfx).argtypes = [P, I, I, I, C.c_longlong, P, P, P, P]
            real = C.c_double if sfx == "_d" else C.c_float
            getattr(lib, "oracle_equilibrium" + sfx).argtypes = [real, real, real, real, P]
            getattr(lib, "oracle_equilibrium" + sfx).restype = None
            getattr(lib, "oracle_les_rate" + sfx).argtypes = [P, real, real]
            getattr(lib, "oracle_les_rate" + sfx).restype = real
            for n in ("oracle_alpha", "oracle_beta", "oracle_init", "oracle_store_dd_region",
                      "oracle_set_dd_region", "oracle_copy_dd_box", "oracle_moments"):
                getattr(lib, n + sfx).restype = None
        lib.oracle_unit.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64, C.c_int]
        lib.oracle_unit.restype = C.c_double
        lib.oracle_num_threads.restype = C.c_int
        _libs[path] = lib
        return lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _i3(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.int32))


def equilibrium(rho, jx, jy, jz, dtype=np.float64):
    """f_eq of the reference expression (lbm_header.h:68-94) for one cell."""
    lib = load()
    out = np.zeros(19, dtype=dtype)
    sfx = "_d" if np.dtype(dtype) == np.float64 else "_f"
    getattr(lib, "oracle_equilibrium" + sfx)(rho, jx, jy, jz, _p(out))
    return out


def les_inv_tau(f, tau0, coeff, dtype=np.float64):
    """1/tau_eff of the oracle's Smagorinsky rate for one cell's 19
    populations, rho and j summed from them; `coeff` is the constant
    18*sqrt(2)*Cs^2 (les_coeff)."""
    lib = load()
    fa = np.ascontiguousarray(np.asarray(f, dtype=dtype))
    sfx = "_d" if np.dtype(dtype) == np.float64 else "_f"
    return float(getattr(lib, "oracle_les_rate" + sfx)(_p(fa), tau0, coeff))


def les_coeff(cs):
    return 18.0 * np.sqrt(2.0) * cs * cs if cs > 0 else 0.0


class OracleDomain:
    """One (sub)domain: dd[19][C], flags[C] (int32), velocity[3][C], density[C]."""

    def __init__(self, dims, bc=(1, 1, 1, 1, 1, 1), *, inv_tau, tau=0.0, smagorinsky_cs=0.0,
                 lid_velocity=0.0, gravitation=(0.0, 0.0, 0.0), forcing=0, dtype=np.float32,
                 store_velocity=True, store_density=True, lib_path=None):
        self.lib = load(lib_path)
        self.dims = tuple(int(d) for d in dims)
        self.C = int(np.prod(self.dims))
        self.dtype = np.dtype(dtype)
        self.sfx = "_d" if self.dtype == np.float64 else "_f"
        self.bc = _i3(bc)
        self.dd = np.zeros((19, self.C), dtype=self.dtype)
        self.flags = np.zeros(self.C, dtype=np.int32)
        self.vel = np.zeros((3, self.C), dtype=self.dtype)
        self.dens = np.zeros(self.C, dtype=self.dtype)
        Params = _ParamsD if self.sfx == "_d" else _ParamsF
        t = self.dtype.type
        self.params = Params(t(inv_tau), t(tau if tau else 1.0 / inv_tau), t(les_coeff(smagorinsky_cs)),
                             t(gravitation[0]), t(gravitation[1]), t(gravitation[2]), t(lid_velocity),
                             int(forcing), int(store_velocity), int(store_density))
        self.step_counter = 0
        self.reset()

    def _fn(self, name):
        return getattr(self.lib, name + self.sfx)

    def reset(self, amplitude=0.0, seed=0, case_id=0, key_origin=(0, 0, 0), key_dims=None):
        """init_kernel (lbm_init.cl:32-237), optionally with the seeded perturbation."""
        kd = _i3(key_dims if key_dims is not None else self.dims)
        ko = _i3(key_origin)
        self._fn("oracle_init")(_p(self.dd), _p(self.flags), _p(self.vel), _p(self.dens), *self.dims,
                                _p(self.bc), float(amplitude), C.c_uint64(seed), C.c_uint64(case_id),
                                _p(ko), _p(kd))
        self.step_counter = 0

    def alpha(self):
        self._fn("oracle_alpha")(_p(self.dd), _p(self.flags), _p(self.vel), _p(self.dens), *self.dims,
                                 C.byref(self.params))

    def beta(self):
        self._fn("oracle_beta")(_p(self.dd), _p(self.flags), _p(self.vel), _p(self.dens), *self.dims,
                                C.byref(self.params))

    def step(self):
        """simulationStep (src/CLbmSolver.hpp:664-676): beta on even counters."""
        if self.step_counter & 1:
            self.alpha()
        else:
            self.beta()
        self.step_counter += 1

    def run(self, n):
        for _ in range(n):
            self.step()

    def set_flags(self, origin, size, value):
        x0, y0, z0 = origin
        sx, sy, sz = size
        f = self.flags.reshape(self.dims[2], self.dims[1], self.dims[0])
        f[z0:z0 + sz, y0:y0 + sy, x0:x0 + sx] = value

    def flags3(self):
        return self.flags.reshape(self.dims[2], self.dims[1], self.dims[0])

    def dd_region(self, origin, size):
        out = np.zeros((19, int(np.prod(size))), dtype=self.dtype)
        self._fn("oracle_store_dd_region")(_p(self.dd), _p(_i3(self.dims)), _p(_i3(origin)), _p(_i3(size)), _p(out))
        return out

    def set_dd_region(self, origin, size, src, norm=None):
        src = np.ascontiguousarray(src, dtype=self.dtype)
        self._fn("oracle_set_dd_region")(_p(self.dd), _p(_i3(self.dims)), _p(_i3(origin)), _p(_i3(size)), _p(src),
                                         _p(_i3(norm)) if norm is not None else None)

    def moments(self):
        """rho, j (float64) per cell from the distributions in the current parity."""
        rho = np.empty(self.C)
        jx = np.empty(self.C)
        jy = np.empty(self.C)
        jz = np.empty(self.C)
        self._fn("oracle_moments")(_p(self.dd), *self.dims, C.c_longlong(self.step_counter), _p(rho), _p(jx),
                                   _p(jy), _p(jz))
        return rho, np.stack([jx, jy, jz])


def copy_box(src: OracleDomain, src_origin, dst: OracleDomain, dst_origin, size, norm=None):
    src.lib  # loaded
    dst._fn("oracle_copy_dd_box")(_p(src.dd), _p(_i3(src.dims)), _p(_i3(src_origin)), _p(dst.dd),
                                  _p(_i3(dst.dims)), _p(_i3(dst_origin)), _p(_i3(size)),
                                  _p(_i3(norm)) if norm is not None else None)


class OracleMulti:
    """A decomposition stepped with the reference's exchange semantics.

    ``subs`` is a list of dicts with keys 'domain' (OracleDomain) and 'comms'
    (list of (dst, send_origin, recv_origin, size, normal))."""

    def __init__(self, subs):
        self.subs = subs
        self.step_counter = 0

    def _partner(self, me, comm):
        dst, _, _, _, normal = comm
        for c in self.subs[dst]["comms"]:
            if c[0] == me and tuple(-n for n in c[4]) == tuple(normal):
                return c
        raise RuntimeError("comm without counterpart")

    def sync(self, after_alpha: bool):
        for axis in range(3):
            for me, sub in enumerate(self.subs):
                for comm in sub["comms"]:
                    dst, send_o, recv_o, size, normal = comm
                    if normal[axis] == 0:
                        continue
                    p = self._partner(me, comm)
                    nb = self.subs[dst]["domain"]
                    if after_alpha:   # syncAlpha: neighbour inner -> my ghost, all 19
                        copy_box(nb, p[1], sub["domain"], recv_o, size, None)
                    else:             # syncBeta: neighbour ghost -> my inner, n.e > 0
                        copy_box(nb, p[2], sub["domain"], send_o, size, normal)

    def step(self):
        """computeNextStep for every subdomain (src/CController.hpp:385-391)."""
        for sub in self.subs:
            sub["domain"].step()
        self.step_counter += 1
        self.sync(after_alpha=(self.step_counter & 1) == 0)

    def run(self, n):
        for _ in range(n):
            self.step()
