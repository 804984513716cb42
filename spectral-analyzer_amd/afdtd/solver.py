"""``AcousticFDTD`` — Python face of the MI355X FDTD engine.

Mirrors the reference's pybind11 class ``_spectral_kernels.AcousticFDTD``
(csrc/bindings/pybind_kernels.cpp:5473-5664, registered at :8924-9022):
same constructor arguments and defaults, same methods, same exception types
and messages.  Every call goes through libafdtd.so (include/afdtd.h); arrays
are copied at the boundary exactly where the reference copies them.

Extension keywords of the constructor (no reference counterpart):
``order`` (4 reference, 2 Yee), ``courant`` (None = reference dt rule),
``boundary`` ('rigid' = rigid edges + PML, 'impedance'), ``refl`` (six
reflection coefficients), ``rho`` / ``bulk`` (per-cell media), ``storage``
('f32', 'bf16', 'f16', 'bfp16': int16 block floating point, 'f16s': FP16 with a
power-of-two scale per 64-cell z-row) and ``kernel`` ('auto', 'generic', 'fused',
'fused2': two steps per sweep).
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import lib

_STORAGE = {"f32": _lib.FDTD_STORE_F32, "bf16": _lib.FDTD_STORE_BF16, "f16": _lib.FDTD_STORE_F16,
            "bfp16": _lib.FDTD_STORE_BFP16, "f16s": _lib.FDTD_STORE_F16S}
_KERNEL = {"auto": _lib.FDTD_KERNEL_AUTO, "generic": _lib.FDTD_KERNEL_GENERIC,
           "fused": _lib.FDTD_KERNEL_FUSED, "fused2": _lib.FDTD_KERNEL_FUSED2}
_BOUNDARY = {"rigid": _lib.FDTD_BOUNDARY_RIGID_PML, "pml": _lib.FDTD_BOUNDARY_RIGID_PML,
             "impedance": _lib.FDTD_BOUNDARY_IMPEDANCE}


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).ravel())


class AcousticFDTD:
    """3-D acoustic FDTD solver with coupled Kirchhoff plate and PML boundaries."""

    def __init__(self, Nx: int, Ny: int, Nz: int, dx: float, c: float = 343.0,
                 rho_air: float = 1.21, cell_type=None, plate_iz: Optional[int] = None,
                 plate_active=None, plate_mass_density: float = 7.0,
                 plate_stiffness_D: float = 0.45, plate_alpha_M: float = 2.0,
                 plate_beta_K: float = 1e-5, n_pml: int = 10, *,
                 order: int = 4, courant: Optional[float] = None, boundary: str = "rigid",
                 refl: Sequence[float] = (1.0,) * 6, rho=None, bulk=None,
                 storage: str = "f32", kernel: str = "auto",
                 tile_y: int = 0, seg_x: int = 0):
        if cell_type is None or plate_iz is None or plate_active is None:
            raise TypeError("AcousticFDTD(): cell_type, plate_iz and plate_active are required")
        self._L = lib()
        self._h = None
        self.Nx, self.Ny, self.Nz = int(Nx), int(Ny), int(Nz)
        ct = np.ascontiguousarray(np.asarray(cell_type, dtype=np.uint8).ravel())
        pa = np.ascontiguousarray(np.asarray(plate_active, dtype=np.uint8).ravel())
        # pybind_kernels.cpp:5491-5494
        if ct.size != self.Nx * self.Ny * self.Nz:
            raise RuntimeError("cell_type: expected Nx*Ny*Nz elements")
        if pa.size != self.Nx * self.Ny:
            raise RuntimeError("plate_active: expected Nx*Ny elements")
        opt = _lib.default_options()
        opt.stencil_order = int(order)
        opt.courant = 0.0 if courant is None else float(courant)
        opt.boundary = _BOUNDARY[boundary]
        for q in range(6):
            opt.refl[q] = float(refl[q])
        self._rho = None if rho is None else _f32(rho)
        self._bulk = None if bulk is None else _f32(bulk)
        opt.rho = None if self._rho is None else self._rho.ctypes.data
        opt.bulk = None if self._bulk is None else self._bulk.ctypes.data
        opt.storage = _STORAGE[storage]
        opt.kernel = _KERNEL[kernel]
        opt.tile_y, opt.seg_x = int(tile_y), int(seg_x)
        h = self._L.fdtd_create_ex(self.Nx, self.Ny, self.Nz, float(dx), float(c),
                                   float(rho_air), _p(ct), int(plate_iz), _p(pa),
                                   float(plate_mass_density), float(plate_stiffness_D),
                                   float(plate_alpha_M), float(plate_beta_K), int(n_pml),
                                   C.byref(opt))
        if not h:
            err = _lib.last_error()
            raise RuntimeError("fdtd_create: allocation failed" + (f" ({err})" if err else ""))
        self._h = h
        self.n_rec = 0
        self._trace_cap = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.fdtd_destroy(h)
            self._h = None

    def close(self):
        self.__del__()

    # ── reference methods (pybind_kernels.cpp:5508-5662) ─────────────────
    def set_bridge_sources(self, cell_indices, weights):
        idx = np.ascontiguousarray(np.asarray(cell_indices, dtype=np.int32).ravel())
        wgt = _f32(weights)
        if idx.size != wgt.size:
            raise RuntimeError("cell_indices and weights must have the same length")
        rc = self._L.fdtd_set_bridge_sources(self._h, int(idx.size), _p(idx), _p(wgt))
        if rc != _lib.FDTD_OK:
            raise RuntimeError(f"fdtd_set_bridge_sources error: {rc}")

    def set_face_fractions(self, vx_frac, vy_frac, vz_frac):
        a = [_f32(v) for v in (vx_frac, vy_frac, vz_frac)]
        nv = self.get_velocity_dims()
        for arr, n in zip(a, nv):
            if arr.size != n:
                raise RuntimeError(f"fdtd_set_face_fractions error: expected {n} elements")
        rc = self._L.fdtd_set_face_fractions(self._h, *[_p(v) for v in a])
        if rc != _lib.FDTD_OK:
            raise RuntimeError(f"fdtd_set_face_fractions error: {rc}")

    def inject_bridge(self, signal_val, signal_ddt, force_scale=1.0):
        rc = self._L.fdtd_inject_bridge(self._h, float(signal_val), float(signal_ddt),
                                        float(force_scale))
        if rc != _lib.FDTD_OK:
            raise RuntimeError(
                "fdtd_inject_bridge is disabled: direct pressure bridge injection "
                "bypasses plate impedance; use AcousticCoEvolver structural coupling")

    def step(self, n_steps: int = 1) -> int:
        rc = self._L.fdtd_step(self._h, int(n_steps))
        if rc == _lib.FDTD_ERR_UNSTABLE:
            raise RuntimeError("AcousticFDTD: pressure diverged (FDTD_ERR_UNSTABLE)")
        if rc == _lib.FDTD_ERR_DEVICE:
            raise RuntimeError(f"AcousticFDTD: device error: {_lib.last_error()}")
        return rc

    def reset(self):
        self._L.fdtd_reset(self._h)

    def get_pressure_field(self) -> np.ndarray:
        out = np.empty((self.Nx, self.Ny, self.Nz), np.float32)
        rc = self._L.fdtd_get_pressure_field(self._h, _p(out), out.size)
        if rc != _lib.FDTD_OK:
            raise RuntimeError(f"fdtd_get_pressure_field error: {rc}")
        return out

    def get_plate_displacement(self) -> np.ndarray:
        out = np.empty((self.Nx, self.Ny), np.float32)
        rc = self._L.fdtd_get_plate_displacement(self._h, _p(out), out.size)
        if rc != _lib.FDTD_OK:
            raise RuntimeError(f"fdtd_get_plate_displacement error: {rc}")
        return out

    def sample_pressure(self, rec_xyz) -> np.ndarray:
        xyz = _f32(rec_xyz)
        n = xyz.size // 3
        out = np.empty(n, np.float32)
        self._L.fdtd_sample_pressure(self._h, n, _p(xyz), _p(out))
        return out

    def sample_velocity(self, rec_xyz):
        xyz = _f32(rec_xyz)
        n = xyz.size // 3
        o = [np.empty(n, np.float32) for _ in range(3)]
        self._L.fdtd_sample_velocity(self._h, n, _p(xyz), *[_p(v) for v in o])
        return tuple(o)

    def get_surface_emission(self, xyz, normals):
        x = _f32(xyz)
        nrm = _f32(normals)
        n = x.size // 3
        P = np.empty(n, np.float32)
        vn = np.empty(n, np.float32)
        self._L.fdtd_get_surface_emission(self._h, n, _p(x), _p(nrm), _p(P), _p(vn))
        return P, vn

    def get_velocity_dims(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        self._L.fdtd_get_velocity_dims(self._h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def _vel(self, fn, n):
        out = np.empty(n, np.float32)
        fn(self._h, _p(out), n)
        return out

    def get_velocity_x(self) -> np.ndarray:
        return self._vel(self._L.fdtd_get_velocity_x, self.get_velocity_dims()[0])

    def get_velocity_y(self) -> np.ndarray:
        return self._vel(self._L.fdtd_get_velocity_y, self.get_velocity_dims()[1])

    def get_velocity_z(self) -> np.ndarray:
        return self._vel(self._L.fdtd_get_velocity_z, self.get_velocity_dims()[2])

    @property
    def dt(self) -> float:
        return float(self._L.fdtd_get_dt(self._h))

    @property
    def step_count(self) -> int:
        return int(self._L.fdtd_get_step_count(self._h))

    # ── C-ABI functions the pybind class does not expose ─────────────────
    def inject_bridge_drive(self, cell_drives, force_scale: float = 1.0) -> int:
        d = _f32(cell_drives)
        return self._L.fdtd_inject_bridge_drive(self._h, _p(d), float(force_scale))

    def inject_plate_force(self, plate_indices, weights, total_force_N: float) -> int:
        i = np.ascontiguousarray(np.asarray(plate_indices, np.int32).ravel())
        w = _f32(weights)
        return self._L.fdtd_inject_plate_force(self._h, int(i.size), _p(i), _p(w),
                                               float(total_force_N))

    def sample_plate_weighted(self, plate_indices, weights) -> float:
        i = np.ascontiguousarray(np.asarray(plate_indices, np.int32).ravel())
        w = _f32(weights)
        out = C.c_float()
        self._L.fdtd_sample_plate_weighted(self._h, int(i.size), _p(i), _p(w), C.byref(out))
        return float(out.value)

    def sample_plate_displacement_batch(self, idx4, wgt4) -> np.ndarray:
        i = np.ascontiguousarray(np.asarray(idx4, np.int32).ravel())
        w = _f32(wgt4)
        out = np.empty(i.size // 4, np.float32)
        self._L.fdtd_sample_plate_displacement_batch(self._h, out.size, _p(i), _p(w), _p(out))
        return out

    def precompute_pressure_samplers(self, rec_xyz):
        xyz = _f32(rec_xyz)
        n = xyz.size // 3
        idx = np.empty(8 * n, np.int32)
        w = np.empty(8 * n, np.float32)
        self._L.fdtd_precompute_pressure_samplers(self._h, n, _p(xyz), _p(idx), _p(w))
        return idx, w

    def precompute_velocity_samplers(self, rec_xyz):
        xyz = _f32(rec_xyz)
        n = xyz.size // 3
        arrs = []
        for _ in range(3):
            arrs += [np.empty(8 * n, np.int32), np.empty(8 * n, np.float32)]
        self._L.fdtd_precompute_velocity_samplers(self._h, n, _p(xyz), *[_p(a) for a in arrs])
        return tuple(arrs)

    def sample_pressure_precomputed(self, idx8, wgt8) -> np.ndarray:
        i = np.ascontiguousarray(np.asarray(idx8, np.int32).ravel())
        w = _f32(wgt8)
        out = np.empty(i.size // 8, np.float32)
        self._L.fdtd_sample_pressure_precomputed(self._h, out.size, _p(i), _p(w), _p(out))
        return out

    def sample_velocity_precomputed(self, xi, xw, yi, yw, zi, zw):
        a = [np.ascontiguousarray(np.asarray(v, np.int32).ravel()) if q % 2 == 0 else _f32(v)
             for q, v in enumerate((xi, xw, yi, yw, zi, zw))]
        n = a[0].size // 8
        o = [np.empty(n, np.float32) for _ in range(3)]
        self._L.fdtd_sample_velocity_precomputed(self._h, n, *[_p(v) for v in a],
                                                 *[_p(v) for v in o])
        return tuple(o)

    # ── extensions ────────────────────────────────────────────────────────
    @property
    def kernel_family(self) -> str:
        return {1: "generic", 2: "fused", 3: "fused2"}.get(
            self._L.fdtd_get_kernel_family(self._h), "?")

    @property
    def steps_per_sweep(self) -> int:
        """Steps one sweep advances: 2 on the two-step sweeps (a room, or an
        order-4 FP32 grid with walls, face fractions or an active plate)."""
        return int(self._L.fdtd_get_sweep_steps(self._h))

    def set_point_sources(self, cell_idx, signals):
        idx = np.ascontiguousarray(np.asarray(cell_idx, np.int32).ravel())
        sig = np.ascontiguousarray(np.asarray(signals, np.float32).reshape(idx.size, -1))
        rc = self._L.fdtd_set_point_sources(self._h, int(idx.size), _p(idx), _p(sig),
                                            int(sig.shape[1]))
        if rc != _lib.FDTD_OK:
            raise ValueError(f"fdtd_set_point_sources error: {rc} {_lib.last_error()}")

    def set_receivers(self, rec_xyz):
        xyz = _f32(rec_xyz)
        self.n_rec = xyz.size // 3
        rc = self._L.fdtd_set_receivers(self._h, self.n_rec, _p(xyz))
        if rc != _lib.FDTD_OK:
            raise RuntimeError(f"fdtd_set_receivers error: {rc}")

    def record_receivers(self, max_steps: int):
        rc = self._L.fdtd_record_receivers(self._h, int(max_steps))
        if rc != _lib.FDTD_OK:
            raise RuntimeError(f"fdtd_record_receivers error: {rc}")
        self._trace_cap = int(max_steps)

    def get_receiver_traces(self) -> np.ndarray:
        out = np.empty(max(self._trace_cap * self.n_rec, 1), np.float32)
        n = C.c_int()
        rc = self._L.fdtd_get_receiver_traces(self._h, _p(out), out.size, C.byref(n))
        if rc != _lib.FDTD_OK:
            raise RuntimeError(f"fdtd_get_receiver_traces error: {rc}")
        return out[: n.value * self.n_rec].reshape(n.value, self.n_rec).copy()

    def energy(self) -> np.ndarray:
        out = np.zeros(2, np.float64)
        rc = self._L.fdtd_energy(self._h, _p(out))
        if rc != _lib.FDTD_OK:
            raise RuntimeError(f"fdtd_energy error: {rc}")
        return out

    def set_state(self, P=None, vx=None, vy=None, vz=None):
        a = [None if v is None else _f32(v) for v in (P, vx, vy, vz)]
        rc = self._L.fdtd_set_state(self._h, *[_p(v) for v in a])
        if rc != _lib.FDTD_OK:
            raise RuntimeError(f"fdtd_set_state error: {rc} {_lib.last_error()}")

    def step_timed(self, n_steps: int):
        """Run n_steps; return (ms_total, ms_per_stencil_launch) from HIP events."""
        tot, sten = C.c_float(), C.c_float()
        rc = self._L.fdtd_step_timed(self._h, int(n_steps), C.byref(tot), C.byref(sten))
        if rc == _lib.FDTD_ERR_UNSTABLE:
            raise RuntimeError("AcousticFDTD: pressure diverged (FDTD_ERR_UNSTABLE)")
        if rc != _lib.FDTD_OK:
            raise RuntimeError(f"fdtd_step_timed error: {rc} {_lib.last_error()}")
        return float(tot.value), float(sten.value)
