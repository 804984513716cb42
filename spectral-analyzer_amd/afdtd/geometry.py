"""Geometry input formats of the uniform solver, computed on the GPU.

Drop-in counterparts of the reference's host-side builders that feed the
solver's C ABI:

* ``voxelise_guitar_body`` (acoustic_fdtd_bridge.py:255-359, the non-AMR
  branch): cell types and the plate mask of a body given by its XY outline
  polygon and rib height, on a grid padded by ``pad_cells`` on every side;
* ``compute_cut_cell_fractions`` (acoustic_fdtd_bridge.py:456-552): open area
  fractions of the faces between a solid and an open cell, from 2 x 2
  sub-points per face, for ``AcousticFDTD.set_face_fractions``.

The grid is derived from the outline's bounding box on the host with the
reference's formulas (it fixes the array shapes); the per-cell and per-face
work runs in ``libafdtd.so`` (``fdtd_voxelise_body``,
``fdtd_cut_cell_fractions``) and equals the reference element for element.
There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from ._lib import FDTD_OK, last_error, lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _outline(outline_pts) -> np.ndarray:
    poly = np.ascontiguousarray(np.asarray(outline_pts, dtype=np.float32)[:, :2])
    if poly.ndim != 2 or poly.shape[0] < 3:
        raise ValueError("outline_pts: expected an (N, 2) polygon with N >= 3")
    return poly


def two_bout_outline(n: int = 96, waist: float = 0.62) -> np.ndarray:
    """A body outline for examples and benchmarks: the union of a lower bout
    (radius 0.18 m, centre y = -0.09 m) and an upper bout (0.13 m, y = 0.10 m)
    seen from the origin, pinched at the waist, as n float32 (x, y)
    vertices, counter-clockwise."""
    pts = []
    for ang in np.linspace(0.0, 2 * np.pi, n, endpoint=False):
        d = np.array([np.cos(ang), np.sin(ang)])
        reach = 0.0
        for c, r in (((0.0, -0.09), 0.18), ((0.0, 0.10), 0.13)):
            b = d @ np.array(c)
            reach = max(reach, b + np.sqrt(b * b - (np.dot(c, c) - r * r)))
        pinch = 1.0 - (1.0 - waist) * np.exp(-(np.sin(ang) / 0.35) ** 2) * abs(np.cos(ang))
        pts.append(reach * pinch * d)
    return np.asarray(pts, np.float32)


def body_grid(outline_pts, body_h: float, dx: float = 0.008, pad_cells: int = 64) -> dict:
    """Grid of ``voxelise_guitar_body``: dimensions, origin, cell centres and
    the back / plate layers (acoustic_fdtd_bridge.py:291-323)."""
    poly = _outline(outline_pts)
    x_min, x_max = float(poly[:, 0].min()), float(poly[:, 0].max())
    y_min, y_max = float(poly[:, 1].min()), float(poly[:, 1].max())
    z_min, z_max = 0.0, body_h
    pad = pad_cells * dx
    gx_min, gx_max = x_min - pad, x_max + pad
    gy_min, gy_max = y_min - pad, y_max + pad
    gz_min, gz_max = z_min - pad, z_max + pad
    Nx = max(4, int(math.ceil((gx_max - gx_min) / dx)))
    Ny = max(4, int(math.ceil((gy_max - gy_min) / dx)))
    Nz = max(4, int(math.ceil((gz_max - gz_min) / dx)))
    xs = gx_min + (np.arange(Nx) + 0.5) * dx
    ys = gy_min + (np.arange(Ny) + 0.5) * dx
    zs = gz_min + (np.arange(Nz) + 0.5) * dx
    return dict(Nx=Nx, Ny=Ny, Nz=Nz, dx=dx, gx_min=gx_min, gy_min=gy_min, gz_min=gz_min,
                z_min=z_min, z_max=z_max, xs=xs, ys=ys, zs=zs,
                iz_back=int(np.argmin(np.abs(zs - z_min))),
                plate_iz=int(np.argmin(np.abs(zs - z_max))))


def voxelise_guitar_body(outline_pts, body_h: float, dx: float = 0.008, pad_cells: int = 64,
                         n_pml: int = 28, *, amr_enabled: bool = False,
                         max_refinement_level: int = 3, importance_policy=None,
                         balance_refinement: bool = True, soundhole=None, progress_cb=None):
    """(cell_type (Nx, Ny, Nz) uint8, plate_active (Nx, Ny) uint8, info).

    Columns whose centre lies inside the outline get the back wall at
    ``iz_back`` and the plate (soundboard) at ``plate_iz``; every other
    column is wall over the body's z range widened by one cell.  ``info``
    holds the grid (``Nx, Ny, Nz, dx, gx_min, gy_min, gz_min, plate_iz,
    iz_back, xs, ys, zs, n_pml``) and ``exterior_mask`` (True outside the
    cavity).  AMR grids (``amr_enabled``) are not part of this engine."""
    if amr_enabled:
        raise NotImplementedError("voxelise_guitar_body: AMR grids are not built by this engine")
    poly = _outline(outline_pts)
    g = body_grid(poly, body_h, dx, pad_cells)
    Nx, Ny, Nz = g["Nx"], g["Ny"], g["Nz"]
    cell_type = np.empty((Nx, Ny, Nz), np.uint8)
    plate_active = np.empty((Nx, Ny), np.uint8)
    rc = lib().fdtd_voxelise_body(_ptr(poly), len(poly), Nx, Ny, Nz, g["gx_min"], g["gy_min"],
                                  g["gz_min"], dx, g["z_min"], g["z_max"], g["iz_back"],
                                  g["plate_iz"], _ptr(cell_type), _ptr(plate_active))
    if rc != FDTD_OK:
        raise RuntimeError(f"voxelise_guitar_body failed (rc={rc}): {last_error()}")
    interior = np.zeros((Nx, Ny, Nz), dtype=bool)
    interior[:, :, g["iz_back"] + 1:g["plate_iz"]] = plate_active.astype(bool)[:, :, None]
    info = {
        "Nx": Nx, "Ny": Ny, "Nz": Nz, "dx": dx,
        "gx_min": g["gx_min"], "gy_min": g["gy_min"], "gz_min": g["gz_min"],
        "plate_iz": g["plate_iz"], "iz_back": g["iz_back"],
        "xs": g["xs"], "ys": g["ys"], "zs": g["zs"], "n_pml": n_pml,
        "exterior_mask": ~interior,
    }
    return cell_type, plate_active, info


def compute_cut_cell_fractions(cell_type, plate_active, outline_pts, info: dict):
    """(vx_frac, vy_frac, vz_frac): flat float32 arrays in the packed
    staggered layouts ((Nx-1)*Ny*Nz, Nx*(Ny-1)*Nz, Nx*Ny*(Nz-1)).

    Faces between two open cells are 1, faces touching a solid cell on both
    sides 0; a face with exactly one solid neighbour gets the share of its
    four sub-points inside the outline and the body's z range
    [gz_min + iz_back dx, gz_min + (plate_iz + 0.5) dx].  The Vz faces just
    below and above the plate layer are 1 under active plate nodes."""
    poly = _outline(outline_pts)
    Nx, Ny, Nz = int(info["Nx"]), int(info["Ny"]), int(info["Nz"])
    ct = np.ascontiguousarray(cell_type, np.uint8).ravel()
    pa = np.ascontiguousarray(plate_active, np.uint8).ravel()
    if ct.size != Nx * Ny * Nz or pa.size != Nx * Ny:
        raise RuntimeError("compute_cut_cell_fractions: cell_type / plate_active do not match info")
    vx = np.empty((Nx - 1) * Ny * Nz, np.float32)
    vy = np.empty(Nx * (Ny - 1) * Nz, np.float32)
    vz = np.empty(Nx * Ny * (Nz - 1), np.float32)
    rc = lib().fdtd_cut_cell_fractions(_ptr(ct), _ptr(pa), _ptr(poly), len(poly), Nx, Ny, Nz,
                                       float(info["gx_min"]), float(info["gy_min"]),
                                       float(info["gz_min"]), float(info["dx"]),
                                       int(info["plate_iz"]), int(info.get("iz_back", 0)),
                                       _ptr(vx), _ptr(vy), _ptr(vz))
    if rc != FDTD_OK:
        raise RuntimeError(f"compute_cut_cell_fractions failed (rc={rc}): {last_error()}")
    return vx, vy, vz


def apply_soundhole_cutout(cell_type, plate_active, info: dict, soundhole) -> None:
    """Open a circular soundhole (cx, cy, r) in the plate layer, in place:
    columns whose centre lies strictly inside the circle lose their plate node
    and their plate-layer cell becomes air (acoustic_fdtd_bridge.py:411-421,
    same float64 expressions)."""
    if soundhole is None:
        return
    cx, cy, r = soundhole
    xs = info["gx_min"] + (np.arange(info["Nx"]) + 0.5) * float(info["dx"])
    ys = info["gy_min"] + (np.arange(info["Ny"]) + 0.5) * float(info["dx"])
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    mask = (X - cx) ** 2 + (Y - cy) ** 2 < r ** 2
    plate_active[mask] = 0
    cell_type[:, :, int(info["plate_iz"])][mask] = 0


def build_bridge_kernel(bridge_positions_xy, info: dict, sigma_m: float = 0.015):
    """(cell_indices int32, weights float32) of the bridge saddle sources
    (acoustic_fdtd_bridge.py:559-599): for each saddle position (x, y), every
    cell of the layer just below the plate whose column centre gives a
    Gaussian weight exp(-r^2 / (2 sigma^2)) of at least 1e-4, in the
    reference's (position, i, j) order and double-precision expression."""
    Nx, Ny, Nz = int(info["Nx"]), int(info["Ny"]), int(info["Nz"])
    xs, ys = info["xs"], info["ys"]
    iz = max(0, int(info["plate_iz"]) - 1)
    sigma2 = sigma_m * sigma_m
    idx, wgt = [], []
    for bx, by in bridge_positions_xy:
        for i in range(Nx):
            dxi = (xs[i] - bx) ** 2
            for j in range(Ny):
                w = math.exp(-(dxi + (ys[j] - by) ** 2) / (2 * sigma2))
                if w < 1e-4:
                    continue
                idx.append(iz + Nz * (j + Ny * i))
                wgt.append(float(w))
    if not idx:
        raise ValueError("No bridge source cells found — check bridge positions overlap the grid")
    return np.array(idx, dtype=np.int32), np.array(wgt, dtype=np.float32)
