"""GPU parity of the AMR pressure backend (include/afdtd_pressure_backend.hpp)
driven through its C++ interface by tests/harness/backend_harness: a
plate-coupled box with bridge drive every step and mic samples after every
step, without and with a room-panel border, equals the oracle's same call
sequence bit for bit (mic traces, final pressure, plate displacement and
the uniform-grid scatter)."""
import os

import numpy as np
import pytest

from oracle.amr_oracle import OracleAMR
from test_backend_cpu import run_harness, write_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n_pml,mode,R", [(0, 0, 0.0), (2, 1, 0.5)])
def test_amr_backend_matches_oracle(tmp_path, n_pml, mode, R):
    from afdtd.amr import AcousticAMRGrid
    from test_amr_oracle_cpu import plated_box
    print(f"backend: plated box n_pml={n_pml} mode={mode}", flush=True)
    grid, kw = plated_box(10, 8, 14, 5)
    if n_pml:
        grid = grid.with_pml_shell(n_pml)
    mics = np.array([[0.0437, 0.0361, 0.0913], [0.0712, 0.0188, 0.0247]], np.float32)
    d = str(tmp_path / "case")
    arrays, sc = write_case(d, grid, kw, n_bridge=4, steps=60, seed=7, n_pml=n_pml,
                            border_mode=mode, border_R=R, mics=mics, uni=(10, 8, 14))
    r = run_harness(d)
    assert r.returncode == 0, r.stdout + r.stderr
    o = OracleAMR(grid.cell_centers, grid.cell_volumes, grid.open_volume_fraction,
                  grid.cell_types, grid.face_cell_neg, grid.face_cell_pos, grid.face_area,
                  grid.face_open_fraction, grid.face_distance, 343.0, 1.21, grid.min_dx)
    o.setup_plate(**kw)
    o.set_bridge_plate_sources(arrays["bridge_idx"], arrays["bridge_wgt"])
    if n_pml:
        o.set_border_condition(mode, 3.0, R, 0.0, n_pml, grid.bounds_min, grid.bounds_max)
    o.setup_mic_samplers(mics, arrays["mic_axis"].reshape(-1, 3), True)
    tr = []
    for s in range(60):
        o.inject_bridge_drive(arrays["drives"][s], 0.01)
        o.step(1)
        tr.append(np.concatenate(o.sample_mics()))
    tr = np.stack(tr).astype(np.float32)
    load = lambda k: np.fromfile(os.path.join(d, k + ".bin"), np.float32)
    assert np.array_equal(load("trace").view(np.uint32), tr.ravel().view(np.uint32))
    assert np.array_equal(load("p").view(np.uint32), o.get_pressure().view(np.uint32))
    assert np.array_equal(load("w").view(np.uint32),
                          o.get_plate_displacement(10 * 8).view(np.uint32))
    uni = o.scatter_pressure_uniform(10, 8, 14, grid.bounds_min, grid.bounds_max)
    assert np.array_equal(load("uni"), uni.ravel())
    assert np.abs(tr).max() > 0
    assert r.stdout.startswith("OK ") and f"dt={o.dt!r}" in r.stdout


def test_amr_backend_idle_silence_and_bridge_drive(tmp_path):
    """The reference's co-evolver checks on the pressure side
    (test_amr_stability.py:291-296, test_amr_coevolver.py:278-338): without
    excitation the backend stays exactly silent; a bridge drive reaches the
    air on both sides of the plate and the mics."""
    from test_amr_oracle_cpu import plated_box
    print("backend: idle silence and bridge drive", flush=True)
    grid, kw = plated_box(10, 8, 14, 5)
    mics = np.array([[0.0437, 0.0361, 0.0913], [0.0512, 0.0388, 0.0215]], np.float32)
    load = lambda d, k: np.fromfile(os.path.join(d, k + ".bin"), np.float32)
    d0 = str(tmp_path / "idle")
    write_case(d0, grid, kw, n_bridge=4, steps=200, mics=mics, uni=(10, 8, 14), drive_scale=0.0)
    r = run_harness(d0)
    assert r.returncode == 0, r.stdout + r.stderr
    assert not np.any(load(d0, "p")) and not np.any(load(d0, "w")) and not np.any(load(d0, "trace"))
    d1 = str(tmp_path / "drive")
    write_case(d1, grid, kw, n_bridge=4, steps=200, mics=mics, uni=(10, 8, 14), drive_scale=1.0)
    r = run_harness(d1)
    assert r.returncode == 0, r.stdout + r.stderr
    p = load(d1, "p").reshape(10, 8, 14)
    assert np.abs(p[:, :, 6:]).max() > 0 and np.abs(p[:, :, :5]).max() > 0   # above and below
    assert np.all(p[:, :, 5] == 0)                                           # the plate layer
    tr = load(d1, "trace").reshape(200, 4, 2)
    assert np.abs(tr[:, 0]).max() > 0 and np.all(np.isfinite(tr))
