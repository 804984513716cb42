"""Instrument descriptors for the co-evolver: the string, pickup and mic
dicts that the reference's scene factory builds
(acoustic_fdtd_bridge.py:1215-1302, ``build_acoustic_coevolver_from_scene``)
and a uniform body dict for ``AcousticCoEvolver``.

String physics follows the reference tables (acoustic_fdtd_bridge.py:1009-1037):
tension from the tuning, scale length and linear mass; axial stiffness from
the gauge (steel, E = 200 GPa); the shared neck mode (65 Hz, 0.18 kg, Q 35).
"""
from __future__ import annotations

import math

import numpy as np

SCALE_LENGTH_M = 0.648
TUNING_HZ = (82.4069, 110.0, 146.832, 195.998, 246.942, 329.628)
GAUGE_IN = (0.046, 0.036, 0.026, 0.017, 0.013, 0.010)
# (tension_N, linear_mass_kgm, damping, stiffness_EI), bass to treble
STRING_PARAMS = ((73.5, 5.35e-3, 6.0, 1.8e-4), (64.5, 3.78e-3, 6.5, 1.2e-4),
                 (61.0, 2.50e-3, 7.0, 8.0e-5), (71.0, 1.75e-3, 8.0, 4.5e-5),
                 (66.0, 1.04e-3, 9.0, 2.8e-5), (63.5, 0.40e-3, 10.0, 1.6e-5))


def string_physical_params(index: int, scale_length_m: float = SCALE_LENGTH_M):
    """(f0, gauge_in, tension_N, mu, damping, EI, EA) — acoustic_fdtd_bridge.py:1027-1037."""
    _t, mu, damping, ei = STRING_PARAMS[index]
    f0 = TUNING_HZ[index]
    gauge = GAUGE_IN[index]
    tension = (2.0 * scale_length_m * f0) ** 2 * mu
    d = gauge * 0.0254
    ea = 2.0e11 * (math.pi * 0.25 * d * d)
    return f0, gauge, tension, mu, damping, ei, ea


def aperture_ring(center, axis, r: float, n: int):
    """n points on a disc of radius r around center, normal to axis
    (acoustic_fdtd_bridge.py:113-130)."""
    ax = np.asarray(axis, np.float64)
    ax = ax / max(float(np.linalg.norm(ax)), 1e-12)
    ref = np.array([1.0, 0.0, 0.0]) if abs(ax[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    u = np.cross(ax, ref)
    u /= max(float(np.linalg.norm(u)), 1e-12)
    v = np.cross(ax, u)
    return [np.asarray(center, np.float32)
            + float(r) * (math.cos(2 * math.pi * k / n) * u
                          + math.sin(2 * math.pi * k / n) * v).astype(np.float32)
            for k in range(n)]


def guitar_defs(body_h: float, bridge_positions, n_strings: int = 6, n_segs: int = 240,
                scale_length_m: float = SCALE_LENGTH_M, neck: bool = True,
                mic_ring: int = 8, mic_aperture_r: float = 0.012):
    """(string_defs, pickup_defs, mic_defs) as the reference factory builds them:
    strings fanned across the bridge saddles at body_h + 10 mm, one bridge
    single-coil 20 mm behind the saddles, a cardioid 75 mm above the top
    aimed down plus a ring of mic_ring mics of radius mic_aperture_r."""
    if bridge_positions:
        y_saddle = float(np.mean([by for _, by in bridge_positions]))
    else:
        y_saddle = -0.070
    y_nut = y_saddle + float(scale_length_m)
    x_span = 0.0088 * max(0, n_strings - 1)
    string_z = body_h + 0.010
    pix = np.linspace(0, 5, n_strings, dtype=int)
    strings = []
    for si in range(n_strings):
        x_str = -x_span / 2.0 + si * (x_span / max(1, n_strings - 1))
        x_nut = x_str * 0.72
        ys = np.linspace(y_nut, y_saddle, n_segs + 1, dtype=np.float32)
        xs = np.linspace(x_nut, x_str, n_segs + 1, dtype=np.float32)
        path = np.column_stack([xs, ys, np.full(n_segs + 1, string_z, np.float32)])
        f0, gauge, tension, mu, damping, ei, ea = string_physical_params(int(pix[si]),
                                                                         scale_length_m)
        d = {"path_xyz": path.astype(np.float32), "tension_N": tension, "linear_mass_kgm": mu,
             "damping": damping, "stiffness_EI": ei, "axial_stiffness_N": ea}
        if neck:
            d.update(neck_freq_hz=65.0, neck_mass_kg=0.18, neck_Q=35.0)
        strings.append(d)
    pickups = [{"type": 0, "pos": np.array([0.0, y_saddle - 0.020, string_z], np.float32),
                "axis": np.array([0.0, 0.0, 1.0], np.float32), "pole_sigma": 0.005,
                "coil_spacing": 0.018, "sensitivity": 1.0, "string_mask": -1}]
    center = np.array([0.0, 0.0, body_h + 0.075], np.float32)
    axis = np.array([0.0, 0.0, -1.0], np.float32)
    base = {"axis": axis, "polar_a": 0.5, "polar_b": 0.5, "gain": 1.0}
    mics = [{"pos": center, **base}]
    if mic_aperture_r > 0 and mic_ring > 0:
        for p in aperture_ring(center, axis, mic_aperture_r, mic_ring):
            mics.append({"pos": p, **base})
    return strings, pickups, mics


def body_dict(cell_type, plate_active, info: dict, bridge_positions, body_h: float,
              bridge_sigma: float = 0.01, n_pml: int = 10, **plate) -> dict:
    """The body dict of the reference constructor (pybind_kernels.cpp:5834-5866)
    for a voxelised body; bridge saddles at (x, y, body_h)."""
    bxyz = np.array([[bx, by, body_h] for bx, by in bridge_positions], np.float32)
    d = {"Nx": info["Nx"], "Ny": info["Ny"], "Nz": info["Nz"], "dx": info["dx"],
         "gx_min": info["gx_min"], "gy_min": info["gy_min"], "gz_min": info["gz_min"],
         "c": 343.0, "rho_air": 1.21,
         "cell_type": np.ascontiguousarray(cell_type, np.uint8).ravel(),
         "plate_iz": info["plate_iz"],
         "plate_active": np.ascontiguousarray(plate_active, np.uint8).ravel(),
         "n_pml": n_pml, "bridge_src_xyz": bxyz, "bridge_sigma": bridge_sigma}
    d.update(plate)
    return d


def _plate_patch(positions, origin, pdx, nx, ny, radius):
    """Plate nodes within `radius` plate cells of each (x, y), with Gaussian
    weights (sigma one plate cell) normalised per position; entries of
    different positions may name the same node."""
    idx, wgt = [], []
    for x, y in positions:
        gi, gj = (x - origin[0]) / pdx, (y - origin[1]) / pdx
        ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
        d2 = (ii - gi) ** 2 + (jj - gj) ** 2
        sel = np.flatnonzero((d2 <= radius * radius).ravel())
        if sel.size == 0:
            sel = np.array([int(np.argmin(d2.ravel()))])
        w = np.exp(-0.5 * d2.ravel()[sel])
        idx.append(sel.astype(np.int32))
        wgt.append((w / w.sum()).astype(np.float32))
    if not idx:
        return np.zeros(0, np.int32), np.zeros(0, np.float32)
    return np.concatenate(idx), np.concatenate(wgt)


def amr_box_desc(nx: int, ny: int, nz: int, kp: int, dx: float, origin, bridge_positions,
                 neck_positions=(), plate_dx: float = None, n_pml: int = 0, border_mode: int = 0,
                 border_R: float = 0.0, gradient_order: int = 2, bridge_radius: float = 1.5,
                 neck_radius: float = 1.5, c: float = 343.0, rho_air: float = 1.21):
    """An AMR body descriptor dict with the keys of the reference's
    ``build_amr_coevolver_descriptor()`` (acoustic_amr.py; read by
    pybind_kernels.cpp:5959-6037): a uniform AMR box of nx x ny x nz cells of
    size dx from `origin`, an optional PML shell, and a Kirchhoff plate in
    layer kp (``AcousticAMRGrid.with_plate_layer``: plate node (i, j) is cell
    (i, j, kp), nodes plate_dx apart, 5 dx by default), with bridge and neck
    plate patches around the given (x, y) positions."""
    from .amr import AcousticAMRGrid
    x0, y0, z0 = (float(v) for v in origin)
    g = AcousticAMRGrid.build_uniform_box(x_min=x0, x_max=x0 + nx * dx, y_min=y0,
                                          y_max=y0 + ny * dx, z_min=z0, z_max=z0 + nz * dx,
                                          dx=dx)
    if n_pml:
        g = g.with_pml_shell(n_pml)
    g, kw = g.with_plate_layer(nx, ny, nz, kp, plate_dx=plate_dx)
    pdx, porg = float(kw["plate_dx"]), [float(v) for v in kw["plate_origin"]]
    bidx, bwgt = _plate_patch(bridge_positions, porg, pdx, nx, ny, bridge_radius)
    nidx, nwgt = _plate_patch(neck_positions, porg, pdx, nx, ny, neck_radius)
    flat = np.asarray(kw["plate_active_flat_idx"], np.int32)
    return {
        "n_cells": g.n_cells, "cell_centers": g.cell_centers, "cell_volumes": g.cell_volumes,
        "open_volume_frac": g.open_volume_fraction, "cell_types": g.cell_types,
        "cell_levels": g.cell_levels, "n_faces": g.n_faces, "face_cell_neg": g.face_cell_neg,
        "face_cell_pos": g.face_cell_pos, "face_area": g.face_area,
        "face_open_frac": g.face_open_fraction, "face_distance": g.face_distance,
        "c": float(c), "rho_air": float(rho_air), "min_dx": float(g.min_dx),
        "bounds_min": g.bounds_min, "bounds_max": g.bounds_max,
        "plate_Nx": nx, "plate_Ny": ny, "plate_dx": pdx,
        "plate_origin": np.array(porg, np.float32), "plate_active": kw["plate_active"],
        "plate_mass_density": kw["plate_rho_h"], "plate_stiffness_D": kw["plate_D"],
        "plate_alpha_M": kw["plate_alpha_M"], "plate_beta_K": kw["plate_beta_K"],
        "n_active_plate": int(flat.size), "plate_active_flat_idx": flat,
        "plate_face_above_starts": kw["face_above_starts"],
        "plate_face_above_idx": kw["face_above_idx"],
        "plate_face_above_wgt": kw["face_above_wgt"],
        "plate_face_below_starts": kw["face_below_starts"],
        "plate_face_below_idx": kw["face_below_idx"],
        "plate_face_below_wgt": kw["face_below_wgt"],
        "plate_cell_above": kw["cell_above"], "plate_cell_below": kw["cell_below"],
        "n_bridge_plate": int(bidx.size), "bridge_plate_idx": bidx, "bridge_plate_wgt": bwgt,
        "n_neck_plate": int(nidx.size), "neck_plate_idx": nidx, "neck_plate_wgt": nwgt,
        "soundhole_cx": 0.0, "soundhole_cy": 0.0, "soundhole_radius": 0.0,
        "border_mode": int(border_mode), "border_sigma_order": 3.0,
        "border_R_reflection": float(border_R), "border_Z_match": 0.0, "n_pml": int(n_pml),
        "gradient_order": int(gradient_order)}
