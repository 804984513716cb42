"""The pressure-backend interface (include/afdtd_pressure_backend.hpp,
reference csrc/include/acoustic_pressure_backend.h) through its C++ test
driver: descriptor validation and its error codes and messages, which are
decided before any device work (no GPU needed)."""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(HERE, "harness", "backend_harness")


def write_case(d, grid, kw, n_bridge=3, steps=10, seed=0, n_pml=0, border_mode=0,
               border_R=0.0, mics=None, uni=(6, 5, 7), drive_scale=0.01):
    """Write one backend case (descriptor arrays, drives, mics) into d."""
    rng = np.random.default_rng(seed)
    flat = np.asarray(kw["plate_active_flat_idx"], np.int32)
    arrays = {
        "cell_centers": np.ascontiguousarray(grid.cell_centers, np.float64),
        "cell_volumes": np.ascontiguousarray(grid.cell_volumes, np.float64),
        "open_volume_frac": np.ascontiguousarray(grid.open_volume_fraction, np.float64),
        "cell_types": np.ascontiguousarray(grid.cell_types, np.uint8),
        "cell_levels": np.zeros(grid.n_cells, np.int16),
        "face_cell_neg": np.ascontiguousarray(grid.face_cell_neg, np.int32),
        "face_cell_pos": np.ascontiguousarray(grid.face_cell_pos, np.int32),
        "face_area": np.ascontiguousarray(grid.face_area, np.float64),
        "face_open_frac": np.ascontiguousarray(grid.face_open_fraction, np.float64),
        "face_distance": np.ascontiguousarray(grid.face_distance, np.float64),
        "bounds": np.concatenate([grid.bounds_min, grid.bounds_max]).astype(np.float64),
        "plate_active": np.asarray(kw["plate_active"], np.uint8),
        "plate_active_flat_idx": flat,
        "face_above_starts": np.asarray(kw["face_above_starts"], np.int32),
        "face_above_idx": np.asarray(kw["face_above_idx"], np.int32),
        "face_above_wgt": np.asarray(kw["face_above_wgt"], np.float32),
        "face_below_starts": np.asarray(kw["face_below_starts"], np.int32),
        "face_below_idx": np.asarray(kw["face_below_idx"], np.int32),
        "face_below_wgt": np.asarray(kw["face_below_wgt"], np.float32),
        "cell_above": np.asarray(kw["cell_above"], np.int32),
        "cell_below": np.asarray(kw["cell_below"], np.int32),
        "bridge_idx": flat[: n_bridge].copy(),
        "bridge_wgt": np.full(n_bridge, 1.0 / max(n_bridge, 1), np.float32),
        "drives": rng.standard_normal((steps, n_bridge)).astype(np.float32),
        "mic_pos": np.asarray(mics, np.float32).ravel(),
        "mic_axis": np.tile(np.float32([0, 0, 1]), len(mics)),
    }
    scal = {"n_cells": grid.n_cells, "n_faces": grid.n_faces, "c": 343.0, "rho_air": 1.21,
            "min_dx": grid.min_dx, "plate_Nx": kw["plate_Nx"], "plate_Ny": kw["plate_Ny"],
            "plate_dx": kw["plate_dx"], "plate_rho_h": kw["plate_rho_h"], "plate_D": kw["plate_D"],
            "plate_alpha_M": kw["plate_alpha_M"], "plate_beta_K": kw["plate_beta_K"],
            "n_active_plate": flat.size, "n_bridge": n_bridge, "border_mode": border_mode,
            "border_sigma_order": 3.0, "border_R": border_R, "n_pml": n_pml, "gradient_order": 2,
            "n_mics": len(mics), "steps": steps, "scale": drive_scale, "uni_Nx": uni[0],
            "uni_Ny": uni[1], "uni_Nz": uni[2]}
    os.makedirs(d, exist_ok=True)
    with open(os.path.join(d, "meta.txt"), "w") as f:
        for k, a in arrays.items():
            a.tofile(os.path.join(d, k + ".bin"))
            f.write(f"a {k} {a.nbytes}\n")
        for k, v in scal.items():
            f.write(f"s {k} {float(v)!r}\n")
    return arrays, scal


def run_harness(d, timeout=120):
    if not os.path.exists(HARNESS):
        subprocess.run(["make"], cwd=os.path.dirname(HARNESS), check=True)
    return subprocess.run([HARNESS, d], capture_output=True, text=True, timeout=timeout)


@pytest.mark.parametrize("spoil,code,msg", [
    ("n_bridge", "DESC_INVALID_REQUIRED_FIELD", "bridge plate mapping is missing or empty"),
    ("plate_D", "DESC_INVALID_PLATE", "plate mass density and stiffness must be > 0"),
    ("min_dx", "DESC_INVALID_REQUIRED_FIELD", "c, rho_air, and min_dx must be > 0"),
    ("n_cells", "DESC_INVALID_REQUIRED_FIELD", "n_cells and n_faces must be > 0"),
])
def test_descriptor_validation(tmp_path, spoil, code, msg):
    from test_amr_oracle_cpu import MIC_POS, plated_box
    grid, kw = plated_box(4, 3, 8, 3)
    d = str(tmp_path / "case")
    write_case(d, grid, kw, mics=MIC_POS[:1] * 0 + 0.02)
    meta = open(os.path.join(d, "meta.txt")).read().splitlines()
    meta = [f"s {spoil} 0.0" if ln.split()[:2] == ["s", spoil] else ln for ln in meta]
    open(os.path.join(d, "meta.txt"), "w").write("\n".join(meta) + "\n")
    r = run_harness(d)
    assert r.returncode == 3, r.stdout + r.stderr
    assert f"CREATE_FAILED {code} {msg}" in r.stdout
