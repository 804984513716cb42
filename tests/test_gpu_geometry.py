"""GPU geometry builders (afdtd.geometry: fdtd_voxelise_body,
fdtd_cut_cell_fractions) against the oracle (oracle/geometry.py), element
for element: cell types, plate mask, exterior mask and the three face
fraction arrays, on a non-convex body outline at two resolutions, both
orientations, a 1000-vertex outline, a grid without padding (plate layer at
the grid edge) and the device-memory entry points."""
import ctypes as C

import numpy as np
import pytest

from oracle import geometry as G

pytestmark = pytest.mark.gpu


def body_outline(n=96):
    from afdtd.geometry import two_bout_outline
    return two_bout_outline(n)


def _check(poly, body_h, dx, pad):
    from afdtd.geometry import compute_cut_cell_fractions, voxelise_guitar_body
    ct, pa, info = voxelise_guitar_body(poly, body_h, dx=dx, pad_cells=pad)
    oct_, opa, og = G.voxelise(poly, body_h, dx, pad)
    print(f"geometry {info['Nx']}x{info['Ny']}x{info['Nz']} dx={dx} pad={pad} n={len(poly)}",
          flush=True)
    for k in ("Nx", "Ny", "Nz", "gx_min", "gy_min", "gz_min", "plate_iz", "iz_back"):
        assert info[k] == og[k], k
    assert np.array_equal(ct, oct_)
    assert np.array_equal(pa, opa)
    assert np.array_equal(info["exterior_mask"], og["exterior_mask"])
    fr = compute_cut_cell_fractions(ct, pa, poly, info)
    ofr = G.face_fractions(oct_, opa, poly, og)
    for name, a, b in zip(("vx", "vy", "vz"), fr, ofr):
        nbad = int(np.count_nonzero(a.view(np.uint32) != b.view(np.uint32)))
        assert nbad == 0, f"{name}: {nbad} of {a.size} fractions differ"
    cut = np.concatenate(fr)
    return ct, pa, info, cut


def test_body_outline_matches_oracle():
    poly = body_outline()
    ct, pa, info, cut = _check(poly, 0.095, 0.008, 16)
    assert pa.sum() > 500 and np.any((cut > 0) & (cut < 1))


def test_fine_grid_and_orientation():
    poly = body_outline(180)
    _check(poly, 0.1, 0.0035, 24)
    _check(poly[::-1].copy(), 0.1, 0.0035, 24)


def test_many_vertices_and_no_padding():
    t = np.linspace(0, 2 * np.pi, 1000, endpoint=False)
    r = 0.2 * (1 + 0.25 * np.cos(3 * t) + 0.05 * np.sin(11 * t))
    poly = np.stack([r * np.cos(t), r * np.sin(t)], 1).astype(np.float32)
    _check(poly, 0.06, 0.011, 4)
    ct, pa, info, _ = _check(poly, 0.075, 0.0191, 0)
    assert info["plate_iz"] == info["Nz"] - 1


def test_device_entry_points_match_host_forms():
    import torch
    from afdtd import lib
    from afdtd.geometry import body_grid, compute_cut_cell_fractions, voxelise_guitar_body
    poly = body_outline()
    ct, pa, info = voxelise_guitar_body(poly, 0.095, dx=0.006, pad_cells=8)
    g = body_grid(poly, 0.095, 0.006, 8)
    Nx, Ny, Nz = g["Nx"], g["Ny"], g["Nz"]
    dpoly = torch.tensor(poly, device="cuda")
    dct = torch.empty(Nx * Ny * Nz, dtype=torch.uint8, device="cuda")
    dpa = torch.empty(Nx * Ny, dtype=torch.uint8, device="cuda")
    fr = [torch.empty(n, dtype=torch.float32, device="cuda")
          for n in ((Nx - 1) * Ny * Nz, Nx * (Ny - 1) * Nz, Nx * Ny * (Nz - 1))]
    torch.cuda.synchronize()
    L = lib()
    p = lambda t: C.c_void_p(t.data_ptr())
    assert L.fdtd_voxelise_body_device(p(dpoly), len(poly), Nx, Ny, Nz, g["gx_min"], g["gy_min"],
                                       g["gz_min"], 0.006, 0.0, 0.095, g["iz_back"], g["plate_iz"],
                                       p(dct), p(dpa), None, None) == 0
    assert L.fdtd_cut_cell_fractions_device(p(dct), p(dpa), p(dpoly), len(poly), Nx, Ny, Nz,
                                            g["gx_min"], g["gy_min"], g["gz_min"], 0.006,
                                            g["plate_iz"], g["iz_back"], p(fr[0]), p(fr[1]),
                                            p(fr[2]), None, None) == 0
    torch.cuda.synchronize()
    assert np.array_equal(dct.cpu().numpy(), ct.ravel())
    assert np.array_equal(dpa.cpu().numpy(), pa.ravel())
    for a, b in zip(fr, compute_cut_cell_fractions(ct, pa, poly, info)):
        assert np.array_equal(a.cpu().numpy(), b)
    # argument checks come before any device work
    assert L.fdtd_voxelise_body_device(None, 3, 4, 4, 4, 0, 0, 0, 0.1, 0, 0.1, 0, 1, None, None,
                                       None, None) == -1
    assert L.fdtd_voxelise_body(p(dpoly), 2, 4, 4, 4, 0, 0, 0, 0.1, 0, 0.1, 0, 1, p(dct),
                                p(dpa)) == -2
