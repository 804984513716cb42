"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py import this module, as the checker.  It wraps libfdtd_oracle.so, a
plain-C restatement of the reference step (csrc/kernels/acoustic_fdtd.cpp),
in float32 (bitwise target) and float64 (accuracy oracle).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libfdtd_oracle.so")
_lib = None


class OrcConfig(C.Structure):
    _fields_ = [
        ("Nx", C.c_int), ("Ny", C.c_int), ("Nz", C.c_int),
        ("dx", C.c_float), ("c", C.c_float), ("rho_air", C.c_float),
        ("plate_iz", C.c_int),
        ("plate_mass", C.c_float), ("plate_D", C.c_float),
        ("plate_alpha_M", C.c_float), ("plate_beta_K", C.c_float),
        ("n_pml", C.c_int), ("order", C.c_int), ("courant", C.c_float),
        ("boundary", C.c_int), ("refl", C.c_float * 6),
    ]


def build(force: bool = False) -> str:
    """Compile libfdtd_oracle.so with its Makefile (gcc, no FMA contraction)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        if L.orc_config_size() != C.sizeof(OrcConfig):
            raise RuntimeError("oracle OrcConfig layout mismatch")
        for suf in ("f32", "f64"):
            R = C.c_float if suf == "f32" else C.c_double
            g = lambda n: getattr(L, f"{n}_{suf}")
            g("orc_create").restype = C.c_void_p
            g("orc_create").argtypes = [C.POINTER(OrcConfig), C.c_void_p, C.c_void_p,
                                        C.c_void_p, C.c_void_p]
            g("orc_destroy").argtypes = [C.c_void_p]
            g("orc_get_dt").restype = C.c_float
            g("orc_get_dt").argtypes = [C.c_void_p]
            for n in ("orc_step", "orc_record_receivers"):
                g(n).argtypes = [C.c_void_p, C.c_int]
            for n in ("orc_get_step_count", "orc_reset"):
                g(n).argtypes = [C.c_void_p]
            g("orc_set_face_fractions").argtypes = [C.c_void_p] + [C.c_void_p] * 3
            g("orc_set_bridge_sources").argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
            g("orc_inject_bridge_drive").argtypes = [C.c_void_p, C.c_void_p, C.c_float]
            g("orc_inject_plate_force").argtypes = [C.c_void_p, C.c_int, C.c_void_p,
                                                    C.c_void_p, C.c_float]
            g("orc_set_point_sources").argtypes = [C.c_void_p, C.c_int, C.c_void_p,
                                                   C.c_void_p, C.c_int]
            g("orc_set_receivers").argtypes = [C.c_void_p, C.c_int, C.c_void_p]
            g("orc_get_traces").argtypes = [C.c_void_p, C.c_void_p, C.c_int]
            g("orc_get_fields").argtypes = [C.c_void_p] + [C.c_void_p] * 4
            g("orc_set_fields").argtypes = [C.c_void_p] + [C.c_void_p] * 4
            g("orc_get_plate").argtypes = [C.c_void_p, C.c_void_p]
            g("orc_get_cell_type").argtypes = [C.c_void_p, C.c_void_p]
            g("orc_sample_pressure").argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
            g("orc_sample_velocity").argtypes = [C.c_void_p, C.c_int, C.c_void_p] + [C.c_void_p] * 3
            g("orc_precompute_pressure_samplers").argtypes = [C.c_void_p, C.c_int, C.c_void_p,
                                                              C.c_void_p, C.c_void_p]
            g("orc_precompute_velocity_samplers").argtypes = [C.c_void_p, C.c_int, C.c_void_p] \
                + [C.c_void_p] * 6
            g("orc_energy").argtypes = [C.c_void_p, C.c_void_p]
            g("orc_sample_plate_weighted").argtypes = [C.c_void_p, C.c_int, C.c_void_p,
                                                       C.c_void_p, C.c_void_p]
            g("orc_sample_plate_batch").argtypes = [C.c_void_p, C.c_int, C.c_void_p,
                                                    C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class OracleFDTD:
    """One oracle instance (precision 'f32' or 'f64').

    Arguments mirror ``AcousticFDTD`` (pybind_kernels.cpp:8950-8966) plus the
    extension options of ``fdtd_create_ex``.
    """

    def __init__(self, Nx, Ny, Nz, dx, c=343.0, rho_air=1.21, cell_type=None,
                 plate_iz=-1, plate_active=None, plate_mass_density=7.0,
                 plate_stiffness_D=0.45, plate_alpha_M=2.0, plate_beta_K=1e-5,
                 n_pml=10, order=4, courant=0.0, boundary=0, refl=(1.0,) * 6,
                 rho=None, bulk=None, precision="f32"):
        self._L = lib()
        self.precision = precision
        self.dtype = np.float32 if precision == "f32" else np.float64
        self.Nx, self.Ny, self.Nz = int(Nx), int(Ny), int(Nz)
        N = self.Nx * self.Ny * self.Nz
        if cell_type is None:
            cell_type = np.zeros(N, np.uint8)
        if plate_active is None:
            plate_active = np.zeros(self.Nx * self.Ny, np.uint8)
        self._ct = np.ascontiguousarray(cell_type, dtype=np.uint8).ravel()
        self._pa = np.ascontiguousarray(plate_active, dtype=np.uint8).ravel()
        self._rho = None if rho is None else np.ascontiguousarray(rho, np.float32).ravel()
        self._bulk = None if bulk is None else np.ascontiguousarray(bulk, np.float32).ravel()
        cf = OrcConfig()
        cf.Nx, cf.Ny, cf.Nz = self.Nx, self.Ny, self.Nz
        cf.dx, cf.c, cf.rho_air = dx, c, rho_air
        cf.plate_iz = plate_iz
        cf.plate_mass, cf.plate_D = plate_mass_density, plate_stiffness_D
        cf.plate_alpha_M, cf.plate_beta_K = plate_alpha_M, plate_beta_K
        cf.n_pml, cf.order, cf.courant, cf.boundary = n_pml, order, courant, boundary
        for q in range(6):
            cf.refl[q] = float(refl[q])
        self._cf = cf
        h = self._f("orc_create")(C.byref(cf), _ptr(self._ct), _ptr(self._pa),
                                  _ptr(self._rho), _ptr(self._bulk))
        if not h:
            raise RuntimeError("oracle create failed")
        self._h = h
        self.n_rec = 0

    def _f(self, name):
        return getattr(self._L, f"{name}_{self.precision}")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._f("orc_destroy")(h)
            self._h = None

    # ── reference API surface ─────────────────────────────────────────
    @property
    def dt(self) -> float:
        return float(self._f("orc_get_dt")(self._h))

    @property
    def step_count(self) -> int:
        return int(self._f("orc_get_step_count")(self._h))

    def step(self, n_steps=1) -> int:
        return int(self._f("orc_step")(self._h, int(n_steps)))

    def reset(self):
        self._f("orc_reset")(self._h)

    def set_face_fractions(self, vx=None, vy=None, vz=None):
        a = [None if v is None else np.ascontiguousarray(v, np.float32) for v in (vx, vy, vz)]
        self._f("orc_set_face_fractions")(self._h, *[_ptr(v) for v in a])

    def set_bridge_sources(self, idx, wgt):
        idx = np.ascontiguousarray(idx, np.int32); wgt = np.ascontiguousarray(wgt, np.float32)
        return self._f("orc_set_bridge_sources")(self._h, len(idx), _ptr(idx), _ptr(wgt))

    def inject_bridge_drive(self, drives, force_scale=1.0):
        d = np.ascontiguousarray(drives, np.float32)
        return self._f("orc_inject_bridge_drive")(self._h, _ptr(d), force_scale)

    def inject_plate_force(self, pidx, wgt, total):
        p = np.ascontiguousarray(pidx, np.int32); w = np.ascontiguousarray(wgt, np.float32)
        return self._f("orc_inject_plate_force")(self._h, len(p), _ptr(p), _ptr(w), total)

    def set_point_sources(self, cell_idx, signals):
        idx = np.ascontiguousarray(cell_idx, np.int32)
        sig = np.ascontiguousarray(signals, np.float32).reshape(len(idx), -1)
        rc = self._f("orc_set_point_sources")(self._h, len(idx), _ptr(idx), _ptr(sig), sig.shape[1])
        if rc:
            raise ValueError("point source on a solid or out-of-range cell")

    def set_receivers(self, rec_xyz):
        xyz = np.ascontiguousarray(rec_xyz, np.float32).reshape(-1, 3)
        self.n_rec = len(xyz)
        self._f("orc_set_receivers")(self._h, self.n_rec, _ptr(xyz))

    def record_receivers(self, max_steps):
        self._f("orc_record_receivers")(self._h, int(max_steps))
        self._cap = int(max_steps)

    def get_receiver_traces(self) -> np.ndarray:
        out = np.zeros(self._cap * max(self.n_rec, 1), np.float64)
        n = self._f("orc_get_traces")(self._h, _ptr(out), out.size)
        return out[: n * self.n_rec].reshape(n, self.n_rec)

    def _fields(self):
        Nx, Ny, Nz = self.Nx, self.Ny, self.Nz
        P = np.zeros(Nx * Ny * Nz, self.dtype)
        vx = np.zeros((Nx - 1) * Ny * Nz, self.dtype)
        vy = np.zeros(Nx * (Ny - 1) * Nz, self.dtype)
        vz = np.zeros(Nx * Ny * (Nz - 1), self.dtype)
        self._f("orc_get_fields")(self._h, _ptr(P), _ptr(vx), _ptr(vy), _ptr(vz))
        return P, vx, vy, vz

    def get_pressure_field(self) -> np.ndarray:
        return self._fields()[0].reshape(self.Nx, self.Ny, self.Nz)

    def get_velocity_x(self):
        return self._fields()[1]

    def get_velocity_y(self):
        return self._fields()[2]

    def get_velocity_z(self):
        return self._fields()[3]

    def get_fields(self):
        return self._fields()

    def set_fields(self, P=None, vx=None, vy=None, vz=None):
        a = [None if v is None else np.ascontiguousarray(v, self.dtype).ravel()
             for v in (P, vx, vy, vz)]
        self._f("orc_set_fields")(self._h, *[_ptr(v) for v in a])

    def get_plate_displacement(self) -> np.ndarray:
        w = np.zeros(self.Nx * self.Ny, self.dtype)
        self._f("orc_get_plate")(self._h, _ptr(w))
        return w.reshape(self.Nx, self.Ny)

    def get_cell_type(self) -> np.ndarray:
        out = np.zeros(self.Nx * self.Ny * self.Nz, np.uint8)
        self._f("orc_get_cell_type")(self._h, _ptr(out))
        return out

    def sample_pressure(self, rec_xyz) -> np.ndarray:
        xyz = np.ascontiguousarray(rec_xyz, np.float32).reshape(-1, 3)
        out = np.zeros(len(xyz), self.dtype)
        self._f("orc_sample_pressure")(self._h, len(xyz), _ptr(xyz), _ptr(out))
        return out

    def sample_velocity(self, rec_xyz):
        xyz = np.ascontiguousarray(rec_xyz, np.float32).reshape(-1, 3)
        o = [np.zeros(len(xyz), self.dtype) for _ in range(3)]
        self._f("orc_sample_velocity")(self._h, len(xyz), _ptr(xyz), *[_ptr(v) for v in o])
        return tuple(o)

    def precompute_pressure_samplers(self, rec_xyz):
        xyz = np.ascontiguousarray(rec_xyz, np.float32).reshape(-1, 3)
        idx = np.zeros(8 * len(xyz), np.int32); w = np.zeros(8 * len(xyz), np.float32)
        self._f("orc_precompute_pressure_samplers")(self._h, len(xyz), _ptr(xyz), _ptr(idx), _ptr(w))
        return idx, w

    def precompute_velocity_samplers(self, rec_xyz):
        xyz = np.ascontiguousarray(rec_xyz, np.float32).reshape(-1, 3)
        n = len(xyz)
        arrs = []
        for _ in range(3):
            arrs += [np.zeros(8 * n, np.int32), np.zeros(8 * n, np.float32)]
        self._f("orc_precompute_velocity_samplers")(self._h, n, _ptr(xyz), *[_ptr(a) for a in arrs])
        return tuple(arrs)

    def sample_plate_weighted(self, plate_indices, weights):
        i = np.ascontiguousarray(plate_indices, np.int32).ravel()
        w = np.ascontiguousarray(weights, np.float32).ravel()
        out = np.zeros(1, self.dtype)
        self._f("orc_sample_plate_weighted")(self._h, i.size, _ptr(i), _ptr(w), _ptr(out))
        return out[0]

    def sample_plate_displacement_batch(self, idx4, wgt4) -> np.ndarray:
        i = np.ascontiguousarray(idx4, np.int32).ravel()
        w = np.ascontiguousarray(wgt4, np.float32).ravel()
        out = np.zeros(i.size // 4, self.dtype)
        self._f("orc_sample_plate_batch")(self._h, out.size, _ptr(i), _ptr(w), _ptr(out))
        return out

    def energy(self) -> np.ndarray:
        out = np.zeros(3, np.float64)
        self._f("orc_energy")(self._h, _ptr(out))
        return out
