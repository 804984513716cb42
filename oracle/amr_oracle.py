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
            "amo_setup_plate": (_i, [_vp, _i, _i, _f, _vp, _vp, _f, _f, _f, _f, _i] + [_vp] * 9),
            "amo_set_bridge_sources": (_i, [_vp, _i, _vp, _vp]),
            "amo_set_neck_sources": (_i, [_vp, _i, _vp, _vp]),
            "amo_inject_bridge_drive": (_i, [_vp, _vp, _i, _f]),
            "amo_inject_plate_force": (_i, [_vp, _i, _vp, _vp, _f]),
            "amo_sample_plate_batch": (_i, [_vp, _i, _vp, _vp, _vp]),
            "amo_sample_plate_weighted": (_i, [_vp, _i, _vp, _vp, _vp]),
            "amo_get_plate": (_i, [_vp, _vp, _i]),
            "amo_setup_mics": (_i, [_vp, _i, _vp, _vp, _i]),
            "amo_sample_mics": (_i, [_vp, _i, _vp, _vp, _vp, _vp]),
            "amo_fornberg_d1": (None, [_vp, _i, _f, _vp]),
            "amo_get_stencil": (_i, [_vp, _vp, _vp]),
        }
        for n, (r, a) in sigs.items():
            getattr(L, n).restype = r
            getattr(L, n).argtypes = a
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def fornberg_d1(x, xi=0.0):
    """First-derivative weights at xi over the nodes x (float32, <= 8 nodes)."""
    xs = np.ascontiguousarray(x, np.float32)
    w = np.zeros(xs.size, np.float32)
    lib().amo_fornberg_d1(_p(xs), xs.size, float(xi), _p(w))
    return w


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

    # -- plate and mic samplers (acoustic_amr.cpp:1412-1792) -------------
    def _rc(self, name, rc):
        if rc != 0:
            raise RuntimeError(f"{name} failed: rc={rc}")
        return rc

    def setup_plate(self, plate_Nx, plate_Ny, plate_dx, plate_origin, plate_active, plate_rho_h,
                    plate_D, plate_alpha_M, plate_beta_K, plate_active_flat_idx,
                    face_above_starts, face_above_idx, face_above_wgt, face_below_starts,
                    face_below_idx, face_below_wgt, cell_above, cell_below):
        a = [np.ascontiguousarray(plate_origin, np.float32),
             np.ascontiguousarray(plate_active, np.uint8),
             np.ascontiguousarray(plate_active_flat_idx, np.int32),
             np.ascontiguousarray(face_above_starts, np.int32),
             np.ascontiguousarray(face_above_idx, np.int32),
             np.ascontiguousarray(face_above_wgt, np.float32),
             np.ascontiguousarray(face_below_starts, np.int32),
             np.ascontiguousarray(face_below_idx, np.int32),
             np.ascontiguousarray(face_below_wgt, np.float32),
             np.ascontiguousarray(cell_above, np.int32),
             np.ascontiguousarray(cell_below, np.int32)]
        self._plate_keep = a
        rc = lib().amo_setup_plate(self._h, int(plate_Nx), int(plate_Ny), float(plate_dx),
                                   _p(a[0]), _p(a[1]), float(plate_rho_h), float(plate_D),
                                   float(plate_alpha_M), float(plate_beta_K), int(a[2].size),
                                   *[_p(x) for x in a[2:]])
        return self._rc("amr_setup_plate", rc)

    def set_bridge_plate_sources(self, idx, wgt):
        i = np.ascontiguousarray(idx, np.int32)
        w = np.ascontiguousarray(wgt, np.float32)
        return self._rc("amr_set_bridge_plate_sources",
                        lib().amo_set_bridge_sources(self._h, i.size, _p(i), _p(w)))

    def set_neck_plate_sources(self, idx, wgt):
        i = np.ascontiguousarray(idx, np.int32)
        w = np.ascontiguousarray(wgt, np.float32)
        return self._rc("amr_set_neck_plate_sources",
                        lib().amo_set_neck_sources(self._h, i.size, _p(i), _p(w)))

    def inject_bridge_drive(self, cell_drives, force_scale=1.0):
        d = np.ascontiguousarray(cell_drives, np.float32)
        return self._rc("amr_inject_bridge_drive",
                        lib().amo_inject_bridge_drive(self._h, _p(d), d.size, float(force_scale)))

    def inject_plate_force(self, plate_idx, wgt, total_force_N):
        i = np.ascontiguousarray(plate_idx, np.int32)
        w = np.ascontiguousarray(wgt, np.float32)
        return self._rc("amr_inject_plate_force",
                        lib().amo_inject_plate_force(self._h, i.size, _p(i), _p(w),
                                                     float(total_force_N)))

    def sample_plate_displacement_batch(self, idx4, wgt4):
        i = np.ascontiguousarray(idx4, np.int32).reshape(-1, 4)
        w = np.ascontiguousarray(wgt4, np.float32).reshape(-1, 4)
        out = np.zeros(len(i), np.float32)
        self._rc("amr_sample_plate_displacement_batch",
                 lib().amo_sample_plate_batch(self._h, len(i), _p(i), _p(w), _p(out)))
        return out

    def sample_plate_weighted(self, plate_idx, wgt):
        i = np.ascontiguousarray(plate_idx, np.int32)
        w = np.ascontiguousarray(wgt, np.float32)
        out = np.zeros(1, np.float32)
        self._rc("amr_sample_plate_weighted",
                 lib().amo_sample_plate_weighted(self._h, i.size, _p(i), _p(w), _p(out)))
        return float(out[0])

    def get_plate_displacement(self, n_nodes):
        out = np.empty(int(n_nodes), np.float32)
        self._rc("amr_get_plate_displacement", lib().amo_get_plate(self._h, _p(out), out.size))
        return out

    def setup_mic_samplers(self, pos_xyz, axis_xyz, need_velocity=True):
        pp = np.ascontiguousarray(pos_xyz, np.float32).reshape(-1, 3)
        ax = np.ascontiguousarray(axis_xyz, np.float32).reshape(-1, 3)
        self._n_mics = len(pp)
        return self._rc("amr_setup_mic_samplers",
                        lib().amo_setup_mics(self._h, len(pp), _p(pp), _p(ax),
                                             int(bool(need_velocity))))

    def sample_mics(self):
        n = getattr(self, "_n_mics", 0)
        o = [np.zeros(n, np.float32) for _ in range(4)]
        self._rc("amr_sample_mics", lib().amo_sample_mics(self._h, n, *[_p(x) for x in o]))
        return tuple(o)

    def get_stencil(self):
        """Order 8: (cells, weights), each (n_faces, 8)."""
        sc = np.zeros((self.n_faces, 8), np.int32)
        sw = np.zeros((self.n_faces, 8), np.float32)
        self._rc("stencil", lib().amo_get_stencil(self._h, _p(sc), _p(sw)))
        return sc, sw
