"""ctypes front end of the AMR CPU oracle (test infrastructure only).

Only tests/ and the ``cpu_baseline`` leg of bench.py import this module, as
the checker.  It wraps libamr_oracle.so, a plain-C restatement of the
reference's order-2 AMR stepper (csrc/kernels/acoustic_amr.cpp), with the
method names of the reference's Python class ``AcousticAMR``
(csrc/bindings/pybind_kernels.cpp:6506-6603, :8877-8920).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libamr_oracle.so")
_lib = None
_vp, _i, _f, _d = C.c_void_p, C.c_int, C.c_float, C.c_double


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            subprocess.run(["make", "-s", "libamr_oracle.so"], cwd=_HERE, check=True)
        L = C.CDLL(_LIB_PATH)
        sigs = {
            "amo_create": (_vp, [_i, _vp, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp,
                                 _d, _d, _d, _i]),
            "amo_destroy": (None, [_vp]),
            "amo_get_dt": (_d, [_vp]),
            "amo_get_step_count": (_i, [_vp]),
            "amo_csr_entries": (_i, [_vp]),
            "amo_reset": (_i, [_vp]),
            "amo_set_state": (_i, [_vp, _vp, _vp]),
            "amo_step": (_i, [_vp, _i]),
            "amo_inject_nearest": (_i, [_vp, _vp, _f]),
            "amo_get_pressure": (_i, [_vp, _vp, _i]),
            "amo_get_velocity": (_i, [_vp, _vp, _i]),
            "amo_get_coeffs": (_i, [_vp, _vp, _vp, _vp]),
            "amo_scatter_uniform": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp]),
            "amo_set_border_condition": (_i, [_vp, _i, _f, _f, _f, _i, _vp, _vp]),
        }
        for n, (r, a) in sigs.items():
            getattr(L, n).restype = r
            getattr(L, n).argtypes = a
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class OracleAMR:
    """CPU restatement of ``_spectral_kernels.AcousticAMR`` (order 2)."""

    def __init__(self, cell_centers, cell_volumes, open_volume_fraction, cell_types,
                 face_cell_neg, face_cell_pos, face_area, face_open_fraction, face_distance,
                 c, rho_air, min_dx, gradient_order=2):
        L = lib()
        self._keep = [np.ascontiguousarray(cell_centers, np.float64),
                      np.ascontiguousarray(cell_volumes, np.float64),
                      np.ascontiguousarray(open_volume_fraction, np.float64),
                      np.ascontiguousarray(cell_types, np.uint8),
                      np.ascontiguousarray(face_cell_neg, np.int32),
                      np.ascontiguousarray(face_cell_pos, np.int32),
                      np.ascontiguousarray(face_area, np.float64),
                      np.ascontiguousarray(face_open_fraction, np.float64),
                      np.ascontiguousarray(face_distance, np.float64)]
        k = self._keep
        self.n_cells = int(k[1].size)
        self.n_faces = int(k[4].size)
        self._h = L.amo_create(self.n_cells, _p(k[0]), _p(k[1]), _p(k[2]), _p(k[3]),
                               self.n_faces, _p(k[4]), _p(k[5]), _p(k[6]), _p(k[7]), _p(k[8]),
                               float(c), float(rho_air), float(min_dx), int(gradient_order))
        if not self._h:
            raise RuntimeError("amr_create failed: invalid AMR topology or allocation failure")

    @classmethod
    def from_grid(cls, grid, c=343.0, rho_air=1.21, gradient_order=2):
        return cls(grid.cell_centers, grid.cell_volumes, grid.open_volume_fraction,
                   grid.cell_types, grid.face_cell_neg, grid.face_cell_pos, grid.face_area,
                   grid.face_open_fraction, grid.face_distance, c, rho_air, grid.min_dx,
                   gradient_order)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().amo_destroy(self._h)
            self._h = None

    @property
    def dt(self):
        return lib().amo_get_dt(self._h)

    @property
    def step_count(self):
        return lib().amo_get_step_count(self._h)

    @property
    def n_csr_entries(self):
        return lib().amo_csr_entries(self._h)

    def step(self, n_steps=1):
        rc = lib().amo_step(self._h, int(n_steps))
        if rc != 0:
            raise RuntimeError(f"amr_step failed: rc={rc}")
        return rc

    def step_rc(self, n_steps=1):
        return lib().amo_step(self._h, int(n_steps))

    def reset(self):
        lib().amo_reset(self._h)

    def set_state(self, p=None, v=None):
        p = None if p is None else np.ascontiguousarray(p, np.float32)
        v = None if v is None else np.ascontiguousarray(v, np.float32)
        lib().amo_set_state(self._h, None if p is None else _p(p), None if v is None else _p(v))

    def inject_pressure_nearest(self, xyz, value):
        x = np.ascontiguousarray(xyz, np.float64).reshape(3)
        rc = lib().amo_inject_nearest(self._h, _p(x), float(value))
        if rc != 0:
            raise RuntimeError(f"amr_inject_pressure_nearest failed: rc={rc}")

    def get_pressure(self):
        out = np.empty(self.n_cells, np.float32)
        lib().amo_get_pressure(self._h, _p(out), self.n_cells)
        return out

    def get_velocity(self):
        out = np.empty(self.n_faces, np.float32)
        lib().amo_get_velocity(self._h, _p(out), self.n_faces)
        return out

    def get_coeffs(self):
        pd = np.empty(self.n_cells, np.float32)
        ps = np.empty(self.n_cells, np.float32)
        vd = np.empty(self.n_faces, np.float32)
        lib().amo_get_coeffs(self._h, _p(pd), _p(ps), _p(vd))
        return pd, ps, vd

    def set_border_condition(self, mode, sigma_order, R_reflection, Z_match, n_pml,
                             bounds_min, bounds_max):
        bmin = np.ascontiguousarray(bounds_min, np.float64)
        bmax = np.ascontiguousarray(bounds_max, np.float64)
        return lib().amo_set_border_condition(self._h, int(mode), float(sigma_order),
                                              float(R_reflection), float(Z_match), int(n_pml),
                                              _p(bmin), _p(bmax))

    def scatter_pressure_uniform(self, Nx, Ny, Nz, bounds_min, bounds_max):
        bmin = np.ascontiguousarray(bounds_min, np.float64)
        bmax = np.ascontiguousarray(bounds_max, np.float64)
        out = np.empty(Nx * Ny * Nz, np.float32)
        rc = lib().amo_scatter_uniform(self._h, Nx, Ny, Nz, _p(bmin), _p(bmax), _p(out))
        if rc != 0:
            raise RuntimeError(f"amr_scatter_pressure_uniform failed: rc={rc}")
        return out.reshape(Nx, Ny, Nz)
