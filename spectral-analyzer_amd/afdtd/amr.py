"""Topology-driven (AMR) acoustic stepper on the GPU: the Python side.

Mirrors the reference's AMR Python surface (gradient orders 2 and 8):

* ``AcousticAMR`` — the pybind class ``_spectral_kernels.AcousticAMR``
  (csrc/bindings/pybind_kernels.cpp:6506-6603, :8877-8920): same constructor
  arguments, methods, properties and ``RuntimeError`` messages, backed by the
  ``amr_*`` C ABI of libafdtd.so (include/afdtd_amr.h);
* ``AcousticAMRGrid`` with ``build_uniform_box`` (acoustic_amr.py:188-311),
  the border-condition specs (acoustic_amr.py:143-170) and
  ``AcousticAMRFDTD`` (acoustic_amr.py:314-375).

The library is the only compute path; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Any

import numpy as np

from . import _lib

BORDER_ANECHOIC = 0       # AMR_BORDER_ANECHOIC (acoustic_amr.h:111)
BORDER_ROOM_PANEL = 1     # AMR_BORDER_ROOM_PANEL (acoustic_amr.h:113)
BORDER_TRANSMISSION = 2   # AMR_BORDER_TRANSMISSION (acoustic_amr.h:115)

_vp, _i, _f, _d = C.c_void_p, C.c_int, C.c_float, C.c_double
_SIGS = {
    "amr_create": (_vp, [_i, _vp, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _d, _d, _d, _i]),
    "amr_get_last_create_error_stage": (C.c_char_p, []),
    "amr_get_last_create_error_message": (C.c_char_p, []),
    "amr_get_create_progress_active": (_i, []),
    "amr_get_create_progress_done_faces": (_i, []),
    "amr_get_create_progress_total_faces": (_i, []),
    "amr_destroy": (None, [_vp]),
    "amr_step": (_i, [_vp, _i]),
    "amr_step_timed": (_i, [_vp, _i, _vp, _vp, _vp]),
    "amr_reset": (_i, [_vp]),
    "amr_set_state": (_i, [_vp, _vp, _vp]),
    "amr_inject_pressure_nearest": (_i, [_vp, _vp, _f]),
    "amr_get_pressure": (_i, [_vp, _vp, _i]),
    "amr_get_velocity": (_i, [_vp, _vp, _i]),
    "amr_get_cell_centers": (_i, [_vp, _vp, _i]),
    "amr_get_dt": (_d, [_vp]),
    "amr_get_step_count": (_i, [_vp]),
    "amr_get_csr_entries": (_i, [_vp]),
    "amr_set_record_layout": (_i, [_vp, _i]),
    "amr_record_mic_samples": (_i, [_vp, _i, _i, _vp, _vp, _vp, _vp, _f, _vp, _vp]),
    "amr_get_record_layout": (_i, [_vp]),
    "amr_scatter_pressure_uniform": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp]),
    "amr_set_border_condition": (_i, [_vp, _i, _f, _f, _f, _i, _vp, _vp]),
    "amr_setup_plate": (_i, [_vp, _i, _i, _f, _vp, _vp, _f, _f, _f, _f, _i] + [_vp] * 9),
    "amr_set_bridge_plate_sources": (_i, [_vp, _i, _vp, _vp]),
    "amr_set_neck_plate_sources": (_i, [_vp, _i, _vp, _vp]),
    "amr_inject_bridge_drive": (_i, [_vp, _vp, _i, _f]),
    "amr_inject_plate_force": (_i, [_vp, _i, _vp, _vp, _f]),
    "amr_sample_plate_displacement_batch": (_i, [_vp, _i, _vp, _vp, _vp]),
    "amr_sample_plate_weighted": (_i, [_vp, _i, _vp, _vp, _vp]),
    "amr_get_plate_displacement": (_i, [_vp, _vp, _i]),
    "amr_setup_mic_samplers": (_i, [_vp, _i, _vp, _vp, _i]),
    "amr_sample_mics": (_i, [_vp, _i, _vp, _vp, _vp, _vp]),
}
_bound = None


def lib():
    global _bound
    if _bound is None:
        L = _lib.lib()
        for n, (r, a) in _SIGS.items():
            getattr(L, n).restype = r
            getattr(L, n).argtypes = a
        _bound = L
    return _bound


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class AcousticAMR:
    """GPU drop-in for ``_spectral_kernels.AcousticAMR`` (gradient orders 2 and 8)."""

    def __init__(self, cell_centers, cell_volumes, open_volume_fraction, cell_types,
                 face_cell_neg, face_cell_pos, face_area, face_open_fraction, face_distance,
                 c, rho_air, min_dx, gradient_order):
        cc = np.ascontiguousarray(cell_centers, np.float64)
        cv = np.ascontiguousarray(cell_volumes, np.float64)
        ov = np.ascontiguousarray(open_volume_fraction, np.float64)
        ct = np.ascontiguousarray(cell_types, np.uint8)
        fn = np.ascontiguousarray(face_cell_neg, np.int32)
        fp = np.ascontiguousarray(face_cell_pos, np.int32)
        fa = np.ascontiguousarray(face_area, np.float64)
        fo = np.ascontiguousarray(face_open_fraction, np.float64)
        fd = np.ascontiguousarray(face_distance, np.float64)
        # shape checks of pybind_kernels.cpp:6536-6543
        if cc.ndim != 2 or cc.shape[1] != 3:
            raise RuntimeError("cell_centers must have shape (n_cells, 3)")
        self.n_cells = int(cc.shape[0])
        if cv.size != self.n_cells or ov.size != self.n_cells or ct.size != self.n_cells:
            raise RuntimeError("cell arrays must have n_cells elements")
        self.n_faces = int(fn.size)
        if (fp.size != self.n_faces or fa.size != self.n_faces or fo.size != self.n_faces
                or fd.size != self.n_faces):
            raise RuntimeError("face arrays must have n_faces elements")
        self._h = lib().amr_create(self.n_cells, _p(cc), _p(cv), _p(ov), _p(ct), self.n_faces,
                                   _p(fn), _p(fp), _p(fa), _p(fo), _p(fd), float(c),
                                   float(rho_air), float(min_dx), int(gradient_order))
        if not self._h:
            raise RuntimeError("amr_create failed: invalid AMR topology or allocation failure")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().amr_destroy(h)
            self._h = None

    # -- reference methods (pybind_kernels.cpp:6564-6601) ------------------
    def step(self, n_steps: int = 1) -> int:
        rc = lib().amr_step(self._h, int(n_steps))
        if rc != 0:
            raise RuntimeError(f"amr_step failed: rc={rc}")
        return rc

    def reset(self) -> None:
        rc = lib().amr_reset(self._h)
        if rc != 0:
            raise RuntimeError(f"amr_reset failed: rc={rc}")

    def inject_pressure_nearest(self, xyz, value: float) -> None:
        x = np.ascontiguousarray(xyz, np.float64)
        if x.size != 3:
            raise RuntimeError("xyz must have 3 elements")
        rc = lib().amr_inject_pressure_nearest(self._h, _p(x), float(value))
        if rc != 0:
            raise RuntimeError(f"amr_inject_pressure_nearest failed: rc={rc}")

    def get_pressure(self) -> np.ndarray:
        out = np.empty(self.n_cells, np.float32)
        rc = lib().amr_get_pressure(self._h, _p(out), self.n_cells)
        if rc != 0:
            raise RuntimeError(f"amr_get_pressure failed: rc={rc}")
        return out

    def get_velocity(self) -> np.ndarray:
        out = np.empty(self.n_faces, np.float32)
        rc = lib().amr_get_velocity(self._h, _p(out), self.n_faces)
        if rc != 0:
            raise RuntimeError(f"amr_get_velocity failed: rc={rc}")
        return out

    @property
    def dt(self) -> float:
        return lib().amr_get_dt(self._h)

    @property
    def step_count(self) -> int:
        return lib().amr_get_step_count(self._h)

    # -- C-ABI functions the pybind class does not wrap --------------------
    def get_cell_centers(self) -> np.ndarray:
        out = np.empty(3 * self.n_cells, np.float32)
        rc = lib().amr_get_cell_centers(self._h, _p(out), out.size)
        if rc != 0:
            raise RuntimeError(f"amr_get_cell_centers failed: rc={rc}")
        return out.reshape(self.n_cells, 3)

    def set_border_condition(self, mode: int, sigma_order: float, R_reflection: float,
                             Z_match: float, n_pml: int, bounds_min, bounds_max) -> int:
        bmin = np.ascontiguousarray(bounds_min, np.float64).reshape(3)
        bmax = np.ascontiguousarray(bounds_max, np.float64).reshape(3)
        rc = lib().amr_set_border_condition(self._h, int(mode), float(sigma_order),
                                            float(R_reflection), float(Z_match), int(n_pml),
                                            _p(bmin), _p(bmax))
        if rc != 0:
            raise RuntimeError(f"amr_set_border_condition failed: rc={rc}")
        return rc

    def scatter_pressure_uniform(self, Nx: int, Ny: int, Nz: int, bounds_min,
                                 bounds_max) -> np.ndarray:
        bmin = np.ascontiguousarray(bounds_min, np.float64).reshape(3)
        bmax = np.ascontiguousarray(bounds_max, np.float64).reshape(3)
        out = np.empty(int(Nx) * int(Ny) * int(Nz), np.float32)
        rc = lib().amr_scatter_pressure_uniform(self._h, int(Nx), int(Ny), int(Nz), _p(bmin),
                                                _p(bmax), _p(out))
        if rc != 0:
            raise RuntimeError(f"amr_scatter_pressure_uniform failed: rc={rc}")
        return out.reshape(int(Nx), int(Ny), int(Nz))

    # -- plate and mic samplers (acoustic_amr.h:165-325) -----------------
    def _rc(self, name, rc):
        if rc != 0:
            raise RuntimeError(f"{name} failed: rc={rc}")
        return rc

    def setup_plate(self, plate_Nx, plate_Ny, plate_dx, plate_origin, plate_active, plate_rho_h,
                    plate_D, plate_alpha_M, plate_beta_K, plate_active_flat_idx,
                    face_above_starts, face_above_idx, face_above_wgt, face_below_starts,
                    face_below_idx, face_below_wgt, cell_above, cell_below) -> int:
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
        rc = lib().amr_setup_plate(self._h, int(plate_Nx), int(plate_Ny), float(plate_dx),
                                   _p(a[0]), _p(a[1]), float(plate_rho_h), float(plate_D),
                                   float(plate_alpha_M), float(plate_beta_K), int(a[2].size),
                                   *[_p(x) for x in a[2:]])
        return self._rc("amr_setup_plate", rc)

    def set_bridge_plate_sources(self, idx, wgt) -> int:
        i = np.ascontiguousarray(idx, np.int32)
        w = np.ascontiguousarray(wgt, np.float32)
        return self._rc("amr_set_bridge_plate_sources",
                        lib().amr_set_bridge_plate_sources(self._h, i.size, _p(i), _p(w)))

    def set_neck_plate_sources(self, idx, wgt) -> int:
        i = np.ascontiguousarray(idx, np.int32)
        w = np.ascontiguousarray(wgt, np.float32)
        return self._rc("amr_set_neck_plate_sources",
                        lib().amr_set_neck_plate_sources(self._h, i.size, _p(i), _p(w)))

    def inject_bridge_drive(self, cell_drives, force_scale: float = 1.0) -> int:
        d = np.ascontiguousarray(cell_drives, np.float32)
        return self._rc("amr_inject_bridge_drive",
                        lib().amr_inject_bridge_drive(self._h, _p(d), d.size, float(force_scale)))

    def inject_plate_force(self, plate_idx, wgt, total_force_N: float) -> int:
        i = np.ascontiguousarray(plate_idx, np.int32)
        w = np.ascontiguousarray(wgt, np.float32)
        return self._rc("amr_inject_plate_force",
                        lib().amr_inject_plate_force(self._h, i.size, _p(i), _p(w),
                                                     float(total_force_N)))

    def sample_plate_displacement_batch(self, idx4, wgt4) -> np.ndarray:
        i = np.ascontiguousarray(idx4, np.int32).reshape(-1, 4)
        w = np.ascontiguousarray(wgt4, np.float32).reshape(-1, 4)
        out = np.zeros(len(i), np.float32)
        self._rc("amr_sample_plate_displacement_batch",
                 lib().amr_sample_plate_displacement_batch(self._h, len(i), _p(i), _p(w),
                                                           _p(out)))
        return out

    def sample_plate_weighted(self, plate_idx, wgt) -> float:
        i = np.ascontiguousarray(plate_idx, np.int32)
        w = np.ascontiguousarray(wgt, np.float32)
        out = np.zeros(1, np.float32)
        self._rc("amr_sample_plate_weighted",
                 lib().amr_sample_plate_weighted(self._h, i.size, _p(i), _p(w), _p(out)))
        return float(out[0])

    def get_plate_displacement(self, n_nodes: int) -> np.ndarray:
        out = np.empty(int(n_nodes), np.float32)
        self._rc("amr_get_plate_displacement",
                 lib().amr_get_plate_displacement(self._h, _p(out), out.size))
        return out

    def setup_mic_samplers(self, pos_xyz, axis_xyz, need_velocity: bool = True) -> int:
        pp = np.ascontiguousarray(pos_xyz, np.float32).reshape(-1, 3)
        ax = np.ascontiguousarray(axis_xyz, np.float32).reshape(-1, 3)
        self._n_mics = len(pp)
        return self._rc("amr_setup_mic_samplers",
                        lib().amr_setup_mic_samplers(self._h, len(pp), _p(pp), _p(ax),
                                                     int(bool(need_velocity))))

    def sample_mics(self):
        n = getattr(self, "_n_mics", 0)
        o = [np.zeros(n, np.float32) for _ in range(4)]
        self._rc("amr_sample_mics", lib().amr_sample_mics(self._h, n, *[_p(x) for x in o]))
        return tuple(o)

    # -- extensions ---------------------------------------------------------
    def set_state(self, p=None, v=None) -> None:
        pa = None if p is None else np.ascontiguousarray(p, np.float32)
        va = None if v is None else np.ascontiguousarray(v, np.float32)
        if pa is not None and pa.size != self.n_cells:
            raise RuntimeError("p must have n_cells elements")
        if va is not None and va.size != self.n_faces:
            raise RuntimeError("v must have n_faces elements")
        rc = lib().amr_set_state(self._h, None if pa is None else _p(pa),
                                 None if va is None else _p(va))
        if rc != 0:
            raise RuntimeError(f"amr_set_state failed: rc={rc}")

    def step_rc(self, n_steps: int = 1) -> int:
        """amr_step's return code without raising (for failure tests)."""
        return lib().amr_step(self._h, int(n_steps))

    def step_timed(self, n_steps: int):
        """(rc, ms_total, ms_face_kernels, ms_cell_kernels) from HIP events."""
        t = (C.c_float * 3)()
        rc = lib().amr_step_timed(self._h, int(n_steps), C.byref(t, 0), C.byref(t, 4),
                                  C.byref(t, 8))
        return rc, float(t[0]), float(t[1]), float(t[2])

    @property
    def n_csr_entries(self) -> int:
        return lib().amr_get_csr_entries(self._h)

    def record_mic_samples(self, n_samples: int, substeps: int, polar_a, polar_b, axis_xyz,
                           gain, rho_c: float):
        """n_samples audio samples of `substeps` steps with every mic's polar
        mix recorded on the device; returns (out (n_samples, n_mics), rc,
        samples done)."""
        m = getattr(self, "_n_mics", 0)
        a = np.ascontiguousarray(polar_a, np.float32).reshape(m)
        b = np.ascontiguousarray(polar_b, np.float32).reshape(m)
        ax = np.ascontiguousarray(axis_xyz, np.float32).reshape(m, 3)
        gn = np.ascontiguousarray(gain, np.float32).reshape(m)
        out = np.zeros((int(n_samples), m), np.float32)
        done = C.c_int(0)
        rc = lib().amr_record_mic_samples(self._h, int(n_samples), int(substeps), _p(a), _p(b),
                                          _p(ax), _p(gn), float(rho_c), _p(out), C.byref(done))
        if rc not in (0, -3):
            raise RuntimeError(f"amr_record_mic_samples failed: rc={rc}")
        return out, rc, int(done.value)

    @property
    def record_layout(self) -> str:
        return {1: "direct", 2: "compact"}.get(lib().amr_get_record_layout(self._h), "?")

    @record_layout.setter
    def record_layout(self, layout: str) -> None:
        code = {"auto": 0, "direct": 1, "compact": 2}[layout]
        if lib().amr_set_record_layout(self._h, code) != 0:
            raise RuntimeError("amr_set_record_layout failed")


def last_create_error() -> tuple[str, str]:
    L = lib()
    return (L.amr_get_last_create_error_stage().decode(),
            L.amr_get_last_create_error_message().decode())


# ---------------------------------------------------------------------------
# Grid descriptor and border specs (acoustic_amr.py:143-311)
# ---------------------------------------------------------------------------

@dataclass
class BorderConditionSpec:
    mode: int = BORDER_ANECHOIC
    sigma_order: float = 3.0
    R_reflection: float = 0.0
    Z_match: float = 0.0


def border_anechoic(sigma_order: float = 3.0) -> BorderConditionSpec:
    return BorderConditionSpec(mode=BORDER_ANECHOIC, sigma_order=sigma_order)


def border_room_panel(R_reflection: float, sigma_order: float = 2.0) -> BorderConditionSpec:
    return BorderConditionSpec(mode=BORDER_ROOM_PANEL, sigma_order=sigma_order,
                               R_reflection=float(R_reflection))


def border_transmission(Z_match: float = 0.0) -> BorderConditionSpec:
    return BorderConditionSpec(mode=BORDER_TRANSMISSION, Z_match=float(Z_match))


@dataclass(frozen=True)
class AcousticAMRGrid:
    base_dx: float
    min_dx: float
    max_refinement_level: int
    bounds_min: np.ndarray
    bounds_max: np.ndarray
    cell_centers: np.ndarray
    cell_half_sizes: np.ndarray
    cell_levels: np.ndarray
    cell_types: np.ndarray
    importance: np.ndarray
    cell_volumes: np.ndarray
    open_volume_fraction: np.ndarray
    face_cell_neg: np.ndarray
    face_cell_pos: np.ndarray
    face_axis: np.ndarray
    face_area: np.ndarray
    face_open_fraction: np.ndarray
    face_distance: np.ndarray
    soundhole: tuple = (0.0, 0.0, 0.0)
    metadata: dict = field(default_factory=dict)

    @property
    def n_cells(self) -> int:
        return int(self.cell_centers.shape[0])

    @property
    def n_faces(self) -> int:
        return int(self.face_cell_neg.shape[0])

    @classmethod
    def build_uniform_box(cls, *, x_min, x_max, y_min, y_max, z_min, z_max,
                          dx) -> "AcousticAMRGrid":
        """All-acoustic uniform box (acoustic_amr.py:213-311): cells in
        ``i*ny*nz + j*nz + k`` order, interior faces only (x faces, then y,
        then z, each in cell order of their neg cell), area dx^2, distance dx."""
        dx = float(dx)
        nx = max(1, round((x_max - x_min) / dx))
        ny = max(1, round((y_max - y_min) / dx))
        nz = max(1, round((z_max - z_min) / dx))
        n = nx * ny * nz
        xs = x_min + dx * (np.arange(nx) + 0.5)
        ys = y_min + dx * (np.arange(ny) + 0.5)
        zs = z_min + dx * (np.arange(nz) + 0.5)
        cc = np.empty((nx, ny, nz, 3), np.float64)
        cc[..., 0] = xs[:, None, None]
        cc[..., 1] = ys[None, :, None]
        cc[..., 2] = zs[None, None, :]
        cid = np.arange(n, dtype=np.int32).reshape(nx, ny, nz)
        negs = [cid[:-1, :, :].ravel(), cid[:, :-1, :].ravel(), cid[:, :, :-1].ravel()]
        steps = (ny * nz, nz, 1)
        neg = np.concatenate(negs).astype(np.int32)
        pos = np.concatenate([a + s for a, s in zip(negs, steps)]).astype(np.int32)
        axis = np.concatenate([np.full(a.size, k, np.uint8) for k, a in enumerate(negs)])
        nf = neg.size
        return cls(
            base_dx=dx, min_dx=dx, max_refinement_level=0,
            bounds_min=np.array([x_min, y_min, z_min], np.float64),
            bounds_max=np.array([x_max, y_max, z_max], np.float64),
            cell_centers=cc.reshape(n, 3),
            cell_half_sizes=np.full(n, dx / 2.0),
            cell_levels=np.zeros(n, np.int16),
            cell_types=np.zeros(n, np.uint8),
            importance=np.zeros(n, np.uint8),
            cell_volumes=np.full(n, dx ** 3),
            open_volume_fraction=np.ones(n),
            face_cell_neg=neg, face_cell_pos=pos, face_axis=axis,
            face_area=np.full(nf, dx ** 2),
            face_open_fraction=np.ones(nf),
            face_distance=np.full(nf, dx),
        )

    def with_pml_shell(self, n_pml: int) -> "AcousticAMRGrid":
        """Copy with the outer ``n_pml`` layers of a uniform box typed PML (3),
        the cells ``amr_set_border_condition`` grades."""
        lo = self.cell_centers - self.bounds_min
        hi = self.bounds_max - self.cell_centers
        d = np.minimum(lo, hi).min(axis=1)
        ct = self.cell_types.copy()
        ct[(d < n_pml * self.min_dx) & (ct == 0)] = 3
        return AcousticAMRGrid(**{**self.__dict__, "cell_types": ct})


    def with_plate_layer(self, nx: int, ny: int, nz: int, kp: int, active=None,
                         plate_dx: float = None, rho_h: float = 7.0 * 0.003, D: float = 0.45,
                         alpha_M: float = 2.0, beta_K: float = 1e-5):
        """A uniform nx x ny x nz box from build_uniform_box with a Kirchhoff
        plate in layer kp: plate node (i, j) is cell (i, j, kp), typed PLATE
        where active; its above face joins it to (i, j, kp+1), its below face
        joins (i, j, kp-1) to it.  Returns (grid, keyword arguments of
        AcousticAMR.setup_plate).  Plate nodes are 5 cells apart by default,
        which keeps the explicit plate update stable at the air time step."""
        act = np.ones((nx, ny), np.uint8) if active is None else np.asarray(active, np.uint8)
        ct = self.cell_types.copy().reshape(nx, ny, nz)
        ct[:, :, kp][act > 0] = 2
        flat = np.flatnonzero(act.ravel()).astype(np.int32)
        i, j = flat // ny, flat % ny
        f0 = (nx - 1) * ny * nz + nx * (ny - 1) * nz          # first z face
        zf = lambda ii, jj, kk: (f0 + (ii * ny + jj) * (nz - 1) + kk).astype(np.int32)
        cell = lambda ii, jj, kk: ((ii * ny + jj) * nz + kk).astype(np.int32)
        n = flat.size
        pdx = 5 * self.min_dx if plate_dx is None else plate_dx
        kw = dict(plate_Nx=nx, plate_Ny=ny, plate_dx=pdx,
                  plate_origin=[float(self.bounds_min[0]), float(self.bounds_min[1]),
                                float(self.bounds_min[2]) + kp * self.min_dx],
                  plate_active=act.ravel(), plate_rho_h=rho_h, plate_D=D,
                  plate_alpha_M=alpha_M, plate_beta_K=beta_K, plate_active_flat_idx=flat,
                  face_above_starts=np.arange(n + 1, dtype=np.int32),
                  face_above_idx=zf(i, j, kp), face_above_wgt=np.ones(n, np.float32),
                  face_below_starts=np.arange(n + 1, dtype=np.int32),
                  face_below_idx=zf(i, j, kp - 1), face_below_wgt=np.ones(n, np.float32),
                  cell_above=cell(i, j, kp + 1), cell_below=cell(i, j, kp - 1))
        return AcousticAMRGrid(**{**self.__dict__, "cell_types": ct.ravel()}), kw


class AcousticAMRFDTD:
    """Handle for the GPU AMR stepper (acoustic_amr.py:314-375)."""

    def __init__(self, grid: AcousticAMRGrid, c: float = 343.0, rho_air: float = 1.21,
                 gradient_order: int = 2):
        if gradient_order not in (2, 8):
            raise ValueError(f"gradient_order must be 2 or 8, got {gradient_order}")
        self.grid = grid
        self.c = float(c)
        self.rho_air = float(rho_air)
        self._c = AcousticAMR(grid.cell_centers, grid.cell_volumes, grid.open_volume_fraction,
                              grid.cell_types, grid.face_cell_neg, grid.face_cell_pos,
                              grid.face_area, grid.face_open_fraction, grid.face_distance,
                              self.c, self.rho_air, float(grid.min_dx), int(gradient_order))

    @property
    def dt(self) -> float:
        return float(self._c.dt)

    @property
    def pressure(self) -> np.ndarray:
        return self._c.get_pressure()

    @property
    def velocity(self) -> np.ndarray:
        return self._c.get_velocity()

    def reset(self) -> None:
        self._c.reset()

    def inject_pressure(self, xyz, value: float) -> None:
        self._c.inject_pressure_nearest(np.asarray(xyz, np.float64).reshape(3), float(value))

    def step(self, n_steps: int = 1) -> None:
        self._c.step(int(n_steps))

    def set_border_condition(self, spec: BorderConditionSpec, n_pml: int) -> None:
        self._c.set_border_condition(spec.mode, spec.sigma_order, spec.R_reflection,
                                     spec.Z_match, n_pml, self.grid.bounds_min,
                                     self.grid.bounds_max)


__all__ = ["AcousticAMR", "AcousticAMRGrid", "AcousticAMRFDTD", "BorderConditionSpec",
           "border_anechoic", "border_room_panel", "border_transmission", "last_create_error",
           "BORDER_ANECHOIC", "BORDER_ROOM_PANEL", "BORDER_TRANSMISSION"]
