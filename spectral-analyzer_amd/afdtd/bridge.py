"""``FDTDCoEvolver``: the reference bridge's uniform-grid orchestrator
(acoustic_fdtd_bridge.py:606-841) over this engine.

The reference class owns an ``AcousticFDTD`` handle built from a body scene
(``from_body_scene``: voxelise the outline, cut the soundhole, construct the
solver, apply the cut-cell face fractions, register the bridge sources) and
exposes the field and receiver API its callers read every audio block.  This
mirror keeps those names, arguments and return shapes; the body comes as its
outline polygon, rib height, bridge positions and soundhole
(``from_outline``) instead of a scene object, whose parsing belongs to the
scene layer above the boundary.  As in the reference, the direct-pressure
driver (``push_driver_sample`` / ``push_driver_block``) is disabled and
raises; ``drive_bridge_block`` is the structural drive this engine adds: per
audio sample the bridge cells' forces go into the plate
(``fdtd_inject_bridge_drive``) and the solver advances ``steps_per_sample``
steps.
"""
from __future__ import annotations

import math

import numpy as np

from .geometry import (apply_soundhole_cutout, build_bridge_kernel,
                       compute_cut_cell_fractions, voxelise_guitar_body)
from .solver import AcousticFDTD


class FDTDCoEvolver:
    """Uniform-grid FDTD of a body beside the ray tracer (acoustic_fdtd_bridge.py:606)."""

    def __init__(self, fdtd: AcousticFDTD, info: dict, bridge_positions,
                 crossover_hz: float = 1500.0, sample_rate: int = 44100,
                 force_scale: float = 1.0):
        self.fdtd = fdtd
        self.info = info
        self.crossover_hz = crossover_hz
        self._sample_rate = sample_rate
        self._force_scale = 1.0               # as the reference (:633): the argument is unused
        self._sig_prev = 0.0
        self._steps_per_sample = max(1, int(math.ceil((1.0 / sample_rate) / fdtd.dt)))
        self._bridge_positions = bridge_positions
        idx, wgt = build_bridge_kernel(bridge_positions, info)
        fdtd.set_bridge_sources(idx, wgt)
        self.bridge_cells = idx

    # ── factory (acoustic_fdtd_bridge.py:647-699) ────────────────────────
    @classmethod
    def from_outline(cls, outline_pts, body_h: float, bridge_positions, soundhole=None,
                     dx: float = 0.004, pad_cells: int = 12, n_pml: int = 10,
                     crossover_hz: float = 1500.0, sample_rate: int = 44100,
                     force_scale: float = 1.0, c: float = 343.0, rho_air: float = 1.21,
                     plate_mass: float = 7.0, plate_stiff: float = 0.45,
                     plate_alpha_M: float = 2.0, plate_beta_K: float = 1e-5,
                     **solver_kwargs) -> "FDTDCoEvolver":
        """``from_body_scene`` for a body given by its outline (N, 2), rib
        height, bridge saddle positions [(x, y), ...] and optional soundhole
        (cx, cy, r); the defaults are the reference factory's."""
        cell_type, plate_active, info = voxelise_guitar_body(
            outline_pts, body_h, dx=dx, pad_cells=pad_cells, n_pml=n_pml)
        apply_soundhole_cutout(cell_type, plate_active, info, soundhole)
        fdtd = AcousticFDTD(info["Nx"], info["Ny"], info["Nz"], dx, c=c, rho_air=rho_air,
                            cell_type=cell_type.ravel(), plate_iz=info["plate_iz"],
                            plate_active=plate_active.ravel(), plate_mass_density=plate_mass,
                            plate_stiffness_D=plate_stiff, plate_alpha_M=plate_alpha_M,
                            plate_beta_K=plate_beta_K, n_pml=n_pml, **solver_kwargs)
        fdtd.set_face_fractions(*compute_cut_cell_fractions(cell_type, plate_active,
                                                            outline_pts, info))
        return cls(fdtd, info, bridge_positions, crossover_hz=crossover_hz,
                   sample_rate=sample_rate, force_scale=force_scale)

    # ── signal injection (acoustic_fdtd_bridge.py:703-720) ───────────────
    def push_driver_sample(self, signal_sum: float) -> None:
        raise RuntimeError(
            "FDTDCoEvolver.push_driver_sample is disabled: it used direct "
            "pressure bridge injection. Use build_acoustic_coevolver_from_scene "
            "and AcousticCoEvolver for structural string-bridge-plate coupling.")

    def push_driver_block(self, signal_block) -> None:
        for s in signal_block:
            self.push_driver_sample(float(s))

    def drive_bridge_block(self, cell_drives, force_scale: float = 1.0) -> None:
        """One row of per-bridge-cell forces (N) per audio sample: each row is
        injected into the plate, then the solver advances steps_per_sample
        steps (raises RuntimeError on divergence, as AcousticFDTD.step)."""
        drives = np.asarray(cell_drives, np.float32).reshape(-1, len(self.bridge_cells))
        for row in drives:
            self.fdtd.inject_bridge_drive(row, force_scale)
            self.fdtd.step(self._steps_per_sample)

    # ── field access (acoustic_fdtd_bridge.py:724-784) ───────────────────
    def get_pressure_field(self) -> np.ndarray:
        return self.fdtd.get_pressure_field()

    def get_plate_displacement(self) -> np.ndarray:
        return self.fdtd.get_plate_displacement()

    def get_volume_slice(self, axis: int = 2, idx: int = -1) -> np.ndarray:
        P = self.get_pressure_field()
        if idx < 0:
            idx = P.shape[axis] // 2
        idx = max(0, min(P.shape[axis] - 1, idx))
        if axis == 0:
            return P[idx, :, :]
        if axis == 1:
            return P[:, idx, :]
        return P[:, :, idx]

    def get_soundboard_pressure(self) -> np.ndarray:
        return self.get_volume_slice(axis=2, idx=self.info["plate_iz"])

    def sample_at_receivers(self, receiver_world_xyz) -> np.ndarray:
        info = self.info
        xyz = np.asarray(receiver_world_xyz)
        gc = np.empty_like(xyz)
        gc[:, 0] = (xyz[:, 0] - info["gx_min"]) / info["dx"]
        gc[:, 1] = (xyz[:, 1] - info["gy_min"]) / info["dx"]
        gc[:, 2] = (xyz[:, 2] - info["gz_min"]) / info["dx"]
        return self.fdtd.sample_pressure(gc.astype(np.float32))

    def reset(self) -> None:
        self.fdtd.reset()
        self._sig_prev = 0.0

    # ── co-evolution blend (acoustic_fdtd_bridge.py:788-821) ─────────────
    def blend_with_ray_transfer(self, H_ray, freq_hz, H_fdtd, crossover=None):
        fc = crossover if crossover is not None else self.crossover_hz
        ratio = (np.asarray(freq_hz) / fc) ** 4
        lo = 1.0 / (1.0 + ratio)
        hi = 1.0 - lo
        return H_fdtd * lo + H_ray * hi

    @property
    def dt(self) -> float:
        return self.fdtd.dt

    @property
    def steps_per_sample(self) -> int:
        return self._steps_per_sample

    def __repr__(self) -> str:
        info = self.info
        return (f"FDTDCoEvolver(grid={info['Nx']}×{info['Ny']}×{info['Nz']}, "
                f"dx={info['dx'] * 1000:.1f}mm, dt={self.dt * 1e6:.1f}µs, "
                f"steps/sample={self._steps_per_sample}, cross={self.crossover_hz:.0f}Hz)")
