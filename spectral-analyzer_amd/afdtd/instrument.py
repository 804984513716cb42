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
