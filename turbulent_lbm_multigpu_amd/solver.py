"""Python handles over the C-ABI: one (sub)domain solver and a solver group.

``LbmSolver`` wraps ``lbm_solver`` (the device side of the reference's
CLbmSolver<T>, src/CLbmSolver.hpp:67-1124); ``LbmGroup`` wraps ``lbm_group``
(several subdomains stepped together with on-device ghost exchange,
src/CController.hpp:265-391).  Arrays cross the boundary as numpy arrays in the
compute type (float32, or float64 for FP64 storage).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _capi
from ._capi import LbmError, SolverConfig, ivec

D3Q19_E = np.array(
    [[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0],
     [1, 1, 0], [-1, -1, 0], [1, -1, 0], [-1, 1, 0],
     [1, 0, 1], [-1, 0, -1], [1, 0, -1], [-1, 0, 1],
     [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1],
     [0, 0, 1], [0, 0, -1], [0, 0, 0]], dtype=np.int64)
"""D3Q19 directions in the reference's order (src/main.cpp:38-66)."""

_COMPONENTS = {_capi.FIELD_DD: 19, _capi.FIELD_VELOCITY: 3, _capi.FIELD_DENSITY: 1, _capi.FIELD_FLAGS: 1,
               _capi.FIELD_DD_RAW: 19}
_RAW_DTYPE = {_capi.STORAGE_F32: np.uint32, _capi.STORAGE_F64: np.uint64, _capi.STORAGE_F16_SHIFTED: np.uint16,
              _capi.STORAGE_BF16_SHIFTED: np.uint16}


def compute_dtype(storage: int):
    return np.float64 if storage == _capi.STORAGE_F64 else np.float32


def les_coefficient(cs: float) -> float:
    """18*sqrt(2)*Cs^2 of the Smagorinsky rate (0 for Cs = 0)."""
    return 18.0 * np.sqrt(2.0) * cs * cs if cs > 0 else 0.0


class LbmSolver:
    """One (sub)domain on the device."""

    def __init__(self, dims: Sequence[int], bc: Sequence[int] = (1, 1, 1, 1, 1, 1), *,
                 inv_tau: float, tau: float = 0.0, smagorinsky_cs: float = 0.0,
                 lid_velocity: float = 0.0, gravitation: Sequence[float] = (0.0, 0.0, 0.0),
                 storage: int = _capi.STORAGE_F32, forcing: int = _capi.FORCING_REFERENCE,
                 store_velocity: bool = False, store_density: bool = False,
                 work_group_size: int = 256, device: int = 0,
                 kernel_variant: int = 0, lds_per_block: int = 0):
        self._lib = _capi.load_library()
        cfg = SolverConfig()
        self._lib.lbm_solver_config_default(C.byref(cfg))
        for a in range(3):
            cfg.dims[a] = int(dims[a])
        for i in range(6):
            cfg.bc[i] = int(bc[i])
        cfg.storage = int(storage)
        cfg.forcing = int(forcing)
        cfg.inv_tau = float(inv_tau)
        cfg.tau = float(tau) if tau else (1.0 / float(inv_tau))
        cfg.smagorinsky_cs = float(smagorinsky_cs)
        for a in range(3):
            cfg.gravitation[a] = float(gravitation[a])
        cfg.lid_velocity = float(lid_velocity)
        cfg.store_velocity = int(bool(store_velocity))
        cfg.store_density = int(bool(store_density))
        cfg.work_group_size = int(work_group_size)
        cfg.device = int(device)
        cfg.kernel_variant = int(kernel_variant)
        cfg.lds_per_block = int(lds_per_block)
        h = C.c_void_p()
        rc = self._lib.lbm_solver_create(C.byref(cfg), C.byref(h))
        if rc != 0:
            raise LbmError(rc, self._lib.lbm_last_error().decode())
        self._h = h
        self.dims = tuple(int(d) for d in dims)
        self.cells = int(np.prod(self.dims))
        self.storage = int(storage)
        self.dtype = compute_dtype(self.storage)
        self.store_velocity = bool(store_velocity)
        self.store_density = bool(store_density)

    # -- plumbing
    @property
    def handle(self):
        return self._h

    def _err(self):
        return self._lib.lbm_solver_error(self._h).decode()

    def _check(self, rc):
        if rc != 0:
            raise LbmError(rc, self._err())

    def close(self):
        if getattr(self, "_h", None):
            rc = self._lib.lbm_solver_destroy(self._h)
            if rc != 0:
                raise LbmError(rc, self._err())
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            if getattr(self, "_h", None) and getattr(self, "_group_owner", None) is None:
                self._lib.lbm_solver_destroy(self._h)
        except Exception:
            pass

    # -- lifecycle / stepping (src/CLbmSolver.hpp:619-683)
    def reset(self):
        self._check(self._lib.lbm_solver_reset(self._h))

    def init_perturbed(self, amplitude: float, seed: int, case_id: int = 0,
                       key_origin: Sequence[int] = (0, 0, 0), key_dims: Optional[Sequence[int]] = None):
        self._check(self._lib.lbm_solver_init_perturbed(
            self._h, float(amplitude), C.c_uint64(seed), C.c_uint64(case_id), ivec(key_origin),
            ivec(key_dims if key_dims is not None else self.dims)))

    def step(self):
        self._check(self._lib.lbm_solver_simulation_step(self._h))

    def step_alpha(self):
        self._check(self._lib.lbm_solver_simulation_step_alpha(self._h))

    def step_beta(self):
        self._check(self._lib.lbm_solver_simulation_step_beta(self._h))

    def run(self, steps: int):
        self._check(self._lib.lbm_solver_run(self._h, int(steps)))

    def wait(self):
        self._check(self._lib.lbm_solver_wait(self._h))

    @property
    def step_counter(self) -> int:
        v = C.c_longlong()
        self._check(self._lib.lbm_solver_get_step_counter(self._h, C.byref(v)))
        return v.value

    @step_counter.setter
    def step_counter(self, value: int):
        self._check(self._lib.lbm_solver_set_step_counter(self._h, int(value)))

    def config(self) -> SolverConfig:
        cfg = SolverConfig()
        self._check(self._lib.lbm_solver_get_config(self._h, C.byref(cfg)))
        return cfg

    def set_lid_velocity(self, lid: float):
        self._check(self._lib.lbm_solver_set_lid_velocity(self._h, float(lid)))

    def set_relaxation(self, inv_tau: float, tau: float, smagorinsky_cs: float):
        self._check(self._lib.lbm_solver_set_relaxation(self._h, float(inv_tau), float(tau), float(smagorinsky_cs)))

    # -- region transfer (src/CLbmSolver.hpp:688-978)
    def _field_dtype(self, field: int):
        if field == _capi.FIELD_FLAGS:
            return np.int32
        if field == _capi.FIELD_DD_RAW:
            return _RAW_DTYPE[self.storage]
        return self.dtype

    def _box(self, origin, size):
        if origin is None:
            return None, None, tuple(self.dims)
        return ivec(origin), ivec(size), tuple(int(s) for s in size)

    def store(self, field: int, origin=None, size=None) -> np.ndarray:
        o, s, shape = self._box(origin, size)
        ncomp = _COMPONENTS[field]
        dt = self._field_dtype(field)
        out = np.empty((ncomp, shape[2], shape[1], shape[0]), dtype=dt)
        self._check(self._lib.lbm_solver_store(self._h, int(field), o, s, out.ctypes.data_as(C.c_void_p)))
        return out if ncomp > 1 else out[0]

    def set(self, field: int, data, origin=None, size=None, norm=None):
        o, s, shape = self._box(origin, size)
        ncomp = _COMPONENTS[field]
        dt = self._field_dtype(field)
        arr = np.ascontiguousarray(np.asarray(data, dtype=dt)).reshape(ncomp, shape[2], shape[1], shape[0])
        self._check(self._lib.lbm_solver_set(self._h, int(field), o, s, arr.ctypes.data_as(C.c_void_p),
                                             ivec(norm)))

    def place_spheres(self, seed: int, fill: float, rmin: float, rmax: float, key_origin=(0, 0, 0),
                      key_dims=None) -> float:
        """Seeded spherical obstacles covering `fill` of the key grid interior;
        returns the covered fraction."""
        a = C.c_double()
        self._check(self._lib.lbm_solver_place_spheres(self._h, C.c_uint64(seed), float(fill), float(rmin),
                                                       float(rmax), ivec(key_origin),
                                                       ivec(key_dims if key_dims is not None else self.dims),
                                                       C.byref(a)))
        return a.value

    def velocity_checksum(self) -> float:
        v = C.c_float()
        self._check(self._lib.lbm_solver_velocity_checksum(self._h, C.byref(v)))
        return v.value

    def mass_energy(self):
        m, k = C.c_double(), C.c_double()
        self._check(self._lib.lbm_solver_mass_energy(self._h, C.byref(m), C.byref(k)))
        return m.value, k.value

    def checkpoint(self) -> dict:
        """Full state for resume: the distributions' stored bit patterns (exact
        for every storage), flags and the step counter (the parity the
        distributions are stored in)."""
        return {"dd_raw": self.store(_capi.FIELD_DD_RAW), "flags": self.store(_capi.FIELD_FLAGS),
                "step_counter": self.step_counter, "storage": self.storage, "dims": self.dims}

    def restore(self, state: dict):
        """Load a checkpoint taken from a solver of the same dims and storage."""
        if tuple(state["dims"]) != self.dims or int(state["storage"]) != self.storage:
            raise ValueError("checkpoint of a different domain or storage")
        self.set(_capi.FIELD_FLAGS, state["flags"])
        self.set(_capi.FIELD_DD_RAW, state["dd_raw"])
        self.step_counter = int(state["step_counter"])

    def moments(self, origin=None, size=None):
        o, s, shape = self._box(origin, size)
        rho = np.empty((shape[2], shape[1], shape[0]), dtype=self.dtype)
        j = np.empty((3, shape[2], shape[1], shape[0]), dtype=self.dtype)
        self._check(self._lib.lbm_solver_moments(self._h, o, s, rho.ctypes.data_as(C.c_void_p),
                                                 j.ctypes.data_as(C.c_void_p)))
        return rho, j


def copy_box(src: LbmSolver, src_origin, dst: LbmSolver, dst_origin, size, norm=None):
    """Device-to-device box copy of distributions (one CComm transfer)."""
    rc = src._lib.lbm_copy_box(src.handle, ivec(src_origin), dst.handle, ivec(dst_origin), ivec(size), ivec(norm))
    if rc != 0:
        raise LbmError(rc, dst._err())


def colocate(solvers: Sequence[LbmSolver]):
    """Moves the solvers' device buffers into one allocation, side by side
    (lbm_solver_colocate): the subdomains of one decomposition step faster
    that way than in separate allocations.  Call before grouping them."""
    if not solvers:
        return
    lib = _capi.load_library()
    arr = (C.c_void_p * len(solvers))(*[s.handle for s in solvers])
    rc = lib.lbm_solver_colocate(arr, len(solvers))
    if rc != 0:
        raise LbmError(rc, lib.lbm_last_error().decode())


class LbmGroup:
    """Subdomains stepped by one fused kernel per half-step plus on-device
    ghost-layer exchange (x faces, then y, then z)."""

    def __init__(self, solvers: Sequence[LbmSolver]):
        self._lib = _capi.load_library()
        arr = (C.c_void_p * len(solvers))(*[s.handle for s in solvers])
        h = C.c_void_p()
        rc = self._lib.lbm_group_create(arr, len(solvers), C.byref(h))
        if rc != 0:
            raise LbmError(rc, self._lib.lbm_last_error().decode())
        self._h = h
        self.solvers = list(solvers)
        for s in self.solvers:
            s._group_owner = self

    @classmethod
    def empty(cls, storage: int = _capi.STORAGE_F32, work_group_size: int = 256, kernel_variant: int = 0,
              device: int = 0) -> "LbmGroup":
        """A group with no members yet (members join with add_member)."""
        g = cls.__new__(cls)
        g._lib = _capi.load_library()
        h = C.c_void_p()
        rc = g._lib.lbm_group_create_empty(int(device), int(storage), int(work_group_size), int(kernel_variant),
                                           C.byref(h))
        if rc != 0:
            raise LbmError(rc, g._lib.lbm_last_error().decode())
        g._h = h
        g.solvers = []
        return g

    def _check(self, rc):
        if rc != 0:
            raise LbmError(rc, self._lib.lbm_group_error(self._h).decode())

    def add_member(self, solver: "LbmSolver") -> int:
        """Join between steps (step counter of the group's parity); returns the member index."""
        idx = C.c_int()
        self._check(self._lib.lbm_group_add_member(self._h, solver.handle, C.byref(idx)))
        self.solvers.append(solver)
        solver._group_owner = self
        return idx.value

    def remove_member(self, index: int) -> "LbmSolver":
        """Leave with the state in place; later members are renumbered."""
        self._check(self._lib.lbm_group_remove_member(self._h, int(index)))
        s = self.solvers.pop(int(index))
        s._group_owner = None
        return s

    def set_member_stop(self, index: int, steps: int):
        """Member `index` freezes after `steps` more half-steps (negative: never)."""
        self._check(self._lib.lbm_group_set_member_stop(self._h, int(index), int(steps)))

    def member_count(self) -> int:
        n = C.c_int()
        self._check(self._lib.lbm_group_member_count(self._h, C.byref(n)))
        return n.value

    def add_comm(self, member: int, dst_member: int, send_origin, recv_origin, size, normal):
        self._check(self._lib.lbm_group_add_comm(self._h, int(member), int(dst_member), ivec(send_origin),
                                                 ivec(recv_origin), ivec(size), ivec(normal)))

    def set_exchange_mode(self, mode: int):
        self._check(self._lib.lbm_group_set_exchange_mode(self._h, int(mode)))

    def set_fused_x(self, mode):
        """Exchange same-sized x faces inside the step kernel (True / 1), by
        the copy launch (False / 0), or per storage (_capi.FUSED_X_AUTO)."""
        self._check(self._lib.lbm_group_set_fused_x(self._h, int(mode)))

    def set_graph(self, enable: bool):
        self._check(self._lib.lbm_group_set_graph(self._h, int(bool(enable))))

    def set_graph_pairs(self, pairs: int):
        """Most beta+alpha pairs one graph launch replays."""
        self._check(self._lib.lbm_group_set_graph_pairs(self._h, int(pairs)))

    def prepare_steps(self, steps: int):
        """Capture the graphs a following step(steps) replays, running nothing."""
        self._check(self._lib.lbm_group_prepare_steps(self._h, int(steps)))

    def launch_counts(self) -> dict:
        """Graph launches and graph-less half-steps of step() so far, graphs captured."""
        a, b, c = C.c_longlong(), C.c_longlong(), C.c_longlong()
        self._check(self._lib.lbm_group_get_launch_counts(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return {"graph_launches": a.value, "eager_half_steps": b.value, "graphs_captured": c.value}

    def step(self, steps: int = 1):
        self._check(self._lib.lbm_group_step(self._h, int(steps)))

    def set_launch(self, work_group_size: int, kernel_variant: int):
        self._check(self._lib.lbm_group_set_launch(self._h, int(work_group_size), int(kernel_variant)))

    def launch(self):
        w, v = C.c_int(), C.c_int()
        self._check(self._lib.lbm_group_get_launch(self._h, C.byref(w), C.byref(v)))
        return w.value, v.value

    def set_alpha_work_group(self, work_group_size: int):
        """Alpha launches at this work-group size (0: the group's); beta keeps the group's."""
        self._check(self._lib.lbm_group_set_alpha_work_group(self._h, int(work_group_size)))

    def alpha_work_group(self) -> int:
        w = C.c_int()
        self._check(self._lib.lbm_group_get_alpha_work_group(self._h, C.byref(w)))
        return w.value

    def autotune(self, pairs: int = 4):
        """Pick the fastest launch shape (work-group size and variant, then
        alpha's own work-group size, then the XCD tile map on or off); the
        timed steps are simulation steps."""
        w, v, ms = C.c_int(), C.c_int(), C.c_double()
        self._check(self._lib.lbm_group_autotune(self._h, int(pairs), C.byref(w), C.byref(v), C.byref(ms)))
        return {"work_group_size": w.value, "kernel_variant": v.value, "alpha_work_group": self.alpha_work_group(),
                "ms_per_step": ms.value}

    def wait(self):
        self._check(self._lib.lbm_group_wait(self._h))

    def sync(self, after_beta: bool):
        """syncAlpha (False) or syncBeta (True) of every member."""
        self._check(self._lib.lbm_group_sync(self._h, int(bool(after_beta))))

    def set_timing(self, enable: bool):
        self._check(self._lib.lbm_group_set_timing(self._h, int(bool(enable))))

    def timing(self):
        a, b = C.c_double(), C.c_double()
        na, nb = C.c_longlong(), C.c_longlong()
        self._check(self._lib.lbm_group_get_timing(self._h, C.byref(a), C.byref(na), C.byref(b), C.byref(nb)))
        out = {"step_kernel_ms": a.value, "step_kernel_launches": na.value,
               "exchange_ms": b.value, "exchange_launches": nb.value}
        be, al, xc = C.c_double(), C.c_double(), C.c_double()
        self._check(self._lib.lbm_group_get_timing_detail(self._h, C.byref(be), C.byref(al), C.byref(xc)))
        out.update(beta_kernel_ms=be.value, alpha_kernel_ms=al.value)
        return out

    def close(self):
        if getattr(self, "_h", None):
            self._check(self._lib.lbm_group_destroy(self._h))
            self._h = None
            for s in self.solvers:
                s._group_owner = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                self._lib.lbm_group_destroy(self._h)
                self._h = None
                for s in self.solvers:
                    s._group_owner = None
        except Exception:
            pass
