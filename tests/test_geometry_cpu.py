"""Pins of the geometry oracle (oracle/geometry.py) without a GPU.

* point-in-polygon against analytic membership: a convex polygon (all
  half-planes), an L-shaped polygon, both orientations;
* the grid on binary-exact numbers, computed by hand;
* cell types, plate mask and exterior mask against per-cell loops with
  analytic convex membership;
* face fractions against per-face loops over the four sub-points with
  analytic convex membership, including the plate-layer rule and a grid
  whose plate layer is the last one (no plate rule).
The polygons are checked to keep every tested point at least 1e-9 away from
their edges, so the analytic answer is unambiguous.
"""
import numpy as np
import pytest

from oracle import geometry as G


def _hexagon(cx=0.013, cy=-0.021, r=0.137, rot=0.3, stretch=1.3):
    t = rot + np.arange(6) * np.pi / 3
    return np.stack([cx + stretch * r * np.cos(t), cy + r * np.sin(t)], 1).astype(np.float32)


def _convex_inside(poly, x, y):
    """Analytic membership of a convex polygon (either orientation)."""
    p = np.asarray(poly, np.float64)
    a, b = p, np.roll(p, -1, axis=0)
    cr = (b[:, 0] - a[:, 0])[:, None] * (np.ravel(y)[None] - a[:, 1:2]) - \
         (b[:, 1] - a[:, 1])[:, None] * (np.ravel(x)[None] - a[:, 0:1])
    area2 = np.sum(a[:, 0] * b[:, 1] - b[:, 0] * a[:, 1])
    inside = np.all(cr > 0, 0) if area2 > 0 else np.all(cr < 0, 0)
    return inside.reshape(np.shape(x))


def _edge_distance(poly, x, y):
    p = np.asarray(poly, np.float64)
    a, b = p, np.roll(p, -1, axis=0)
    P = np.stack([np.ravel(x), np.ravel(y)], 1)[:, None, :]
    ab = (b - a)[None]
    t = np.clip(np.sum((P - a[None]) * ab, 2) / np.sum(ab * ab, 2), 0, 1)
    d = np.linalg.norm(P - (a[None] + t[..., None] * ab), axis=2)
    return d.min(1)


def test_point_in_polygon_convex_both_orientations():
    poly = _hexagon()
    rng = np.random.default_rng(0)
    x = rng.uniform(-0.3, 0.3, 4000)
    y = rng.uniform(-0.3, 0.3, 4000)
    keep = _edge_distance(poly, x, y) > 1e-9
    x, y = x[keep], y[keep]
    want = _convex_inside(poly, x, y)
    assert 500 < want.sum() < len(x) - 500
    assert np.array_equal(G.point_in_polygon(x, y, poly), want)
    assert np.array_equal(G.point_in_polygon(x, y, poly[::-1].copy()), want)


def test_point_in_polygon_l_shape():
    poly = np.array([[0, 0], [2, 0], [2, 1], [1, 1], [1, 2], [0, 2]], np.float32)
    rng = np.random.default_rng(1)
    x = rng.uniform(-0.5, 2.5, 5000)
    y = rng.uniform(-0.5, 2.5, 5000)
    keep = _edge_distance(poly, x, y) > 1e-9
    x, y = x[keep], y[keep]
    want = ((0 < x) & (x < 2) & (0 < y) & (y < 1)) | ((0 < x) & (x < 1) & (0 < y) & (y < 2))
    assert np.array_equal(G.point_in_polygon(x, y, poly), want)
    assert np.array_equal(G.point_in_polygon(x, y, poly[::-1].copy()), want)


def test_grid_binary_exact():
    sq = np.array([[-0.5, -0.5], [0.5, -0.5], [0.5, 0.5], [-0.5, 0.5]], np.float32)
    g = G.grid(sq, 0.25, 0.125, 2)
    # pad 0.25: x, y from -0.75 to 0.75 (12 cells), z from -0.25 to 0.5 (6 cells)
    assert (g["Nx"], g["Ny"], g["Nz"]) == (12, 12, 6)
    assert (g["gx_min"], g["gy_min"], g["gz_min"]) == (-0.75, -0.75, -0.25)
    assert np.array_equal(g["zs"], [-0.1875, -0.0625, 0.0625, 0.1875, 0.3125, 0.4375])
    # ties |zs - 0| and |zs - 0.25| resolve to the first index
    assert (g["iz_back"], g["plate_iz"]) == (1, 3)
    assert G.grid(sq, 0.25, 0.125, 0)["Nx"] == 8          # at least 4 cells, 1.0 / 0.125


def _brute_cells(poly, g):
    Nx, Ny, Nz, dx = g["Nx"], g["Ny"], g["Nz"], g["dx"]
    ct = np.zeros((Nx, Ny, Nz), np.uint8)
    pa = np.zeros((Nx, Ny), np.uint8)
    ext = np.ones((Nx, Ny, Nz), bool)
    for i in range(Nx):
        for j in range(Ny):
            X, Y = g["gx_min"] + (i + 0.5) * dx, g["gy_min"] + (j + 0.5) * dx
            assert _edge_distance(poly, X, Y)[0] > 1e-9
            if _convex_inside(poly, np.array(X), np.array(Y)):
                pa[i, j] = 1
                ct[i, j, g["iz_back"]] = G.WALL
                ct[i, j, g["plate_iz"]] = G.PLATE
                for k in range(g["iz_back"] + 1, g["plate_iz"]):
                    ext[i, j, k] = False
            else:
                for k in range(Nz):
                    z = g["gz_min"] + (k + 0.5) * dx
                    if -dx < z < g["z_max"] + dx:
                        ct[i, j, k] = G.WALL
    return ct, pa, ext


def _brute_fractions(poly, ct, pa, g):
    Nx, Ny, Nz, dx = g["Nx"], g["Ny"], g["Nz"], g["dx"]
    org = (g["gx_min"], g["gy_min"], g["gz_min"])
    zlo = org[2] + g["iz_back"] * dx
    zhi = org[2] + (g["plate_iz"] + 0.5) * dx
    solid = (ct == G.WALL) | (ct == G.PLATE)
    out = []
    for ax in range(3):
        shape = [Nx, Ny, Nz]
        shape[ax] -= 1
        f = np.zeros(shape, np.float32)
        for i in range(shape[0]):
            for j in range(shape[1]):
                for k in range(shape[2]):
                    n = [i, j, k]
                    n[ax] += 1
                    sL, sR = solid[i, j, k], solid[tuple(n)]
                    if ax == 2 and 0 < g["plate_iz"] < Nz - 1 and k in (g["plate_iz"] - 1, g["plate_iz"]):
                        f[i, j, k] = 1.0 if (pa[i, j] or not (sL or sR)) else 0.0
                        continue
                    if sL == sR:
                        f[i, j, k] = 0.0 if sL else 1.0
                        continue
                    hits = 0
                    t = [q for q in range(3) if q != ax]
                    for a in (0.25, 0.75):
                        for b in (0.25, 0.75):
                            c = [0, 0, 0]
                            c[ax] = [i, j, k][ax] + 1
                            c[t[0]] = [i, j, k][t[0]] + a
                            c[t[1]] = [i, j, k][t[1]] + b
                            x, y, z = (o + v * dx for o, v in zip(org, c))
                            if zlo <= z <= zhi:
                                assert _edge_distance(poly, x, y)[0] > 1e-9
                                hits += bool(_convex_inside(poly, np.array(x), np.array(y)))
                    f[i, j, k] = hits / 4.0
        out.append(f.ravel())
    return out


@pytest.mark.parametrize("body_h,dx,pad", [(0.053, 0.0191, 3), (0.004, 0.0173, 2)])
def test_voxelise_and_fractions_against_per_cell_loops(body_h, dx, pad):
    poly = _hexagon()
    ct, pa, g = G.voxelise(poly, body_h, dx, pad)
    bct, bpa, bext = _brute_cells(poly, g)
    assert np.array_equal(ct, bct)
    assert np.array_equal(pa, bpa)
    assert np.array_equal(g["exterior_mask"], bext)
    assert pa.sum() > 20 and (ct == G.WALL).sum() > 0 and (ct == G.PLATE).sum() == pa.sum()
    fr = G.face_fractions(ct, pa, poly, g)
    bfr = _brute_fractions(poly, ct, pa, g)
    for a, b in zip(fr, bfr):
        assert np.array_equal(a, b)
    partial = np.concatenate(fr)
    assert np.any((partial > 0) & (partial < 1))        # cut faces occur


def test_plate_rule_needs_an_interior_plate_layer():
    """With no padding in z the plate layer can be the last one: then the
    plate-layer rule does not apply and those faces follow the cut rule."""
    poly = _hexagon()
    ct, pa, g = G.voxelise(poly, 0.075, 0.0191, 0)
    assert g["plate_iz"] == g["Nz"] - 1
    fr = G.face_fractions(ct, pa, poly, g)
    bfr = _brute_fractions(poly, ct, pa, g)
    for a, b in zip(fr, bfr):
        assert np.array_equal(a, b)


def test_bridge_kernel_and_soundhole_brute_force():
    """build_bridge_kernel (acoustic_fdtd_bridge.py:559-599) against a
    per-position, per-column loop; apply_soundhole_cutout (:411-421) against
    a per-column loop of the circle test."""
    import math as m
    from afdtd.geometry import apply_soundhole_cutout, build_bridge_kernel, two_bout_outline
    from oracle import geometry as og
    poly = two_bout_outline(64)
    ct, pa, g = og.voxelise(poly, 0.1, 0.008, 6)
    bridges = [(0.01, -0.1), (-0.03, -0.12)]
    idx, wgt = build_bridge_kernel(bridges, g, sigma_m=0.02)
    ei, ew = [], []
    iz = g["plate_iz"] - 1
    for bx, by in bridges:
        for i in range(g["Nx"]):
            for j in range(g["Ny"]):
                x = g["gx_min"] + (i + 0.5) * g["dx"]
                y = g["gy_min"] + (j + 0.5) * g["dx"]
                w = m.exp(-((x - bx) ** 2 + (y - by) ** 2) / (2 * 0.02 ** 2))
                if w >= 1e-4:
                    ei.append(iz + g["Nz"] * (j + g["Ny"] * i))
                    ew.append(w)
    assert np.array_equal(idx, np.array(ei, np.int32))
    np.testing.assert_allclose(wgt, np.array(ew, np.float32), rtol=1e-6)
    assert np.all(idx % g["Nz"] == iz)
    ct2, pa2 = ct.copy(), pa.copy()
    apply_soundhole_cutout(ct2, pa2, g, (0.0, 0.05, 0.045))
    for i in range(g["Nx"]):
        for j in range(g["Ny"]):
            x = g["gx_min"] + (i + 0.5) * g["dx"]
            y = g["gy_min"] + (j + 0.5) * g["dx"]
            hole = (x - 0.0) ** 2 + (y - 0.05) ** 2 < 0.045 ** 2
            assert pa2[i, j] == (0 if hole else pa[i, j])
            assert ct2[i, j, g["plate_iz"]] == (0 if hole else ct[i, j, g["plate_iz"]])
    assert (pa2 != pa).sum() > 10
    assert np.array_equal(np.delete(ct2, g["plate_iz"], 2), np.delete(ct, g["plate_iz"], 2))
