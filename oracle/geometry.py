"""Oracle (test infrastructure only): numpy restatement of the reference's
geometry builders that feed the uniform solver.  Imported by tests/ and
bench.py's CPU baseline as the checker of the GPU geometry kernels, never by
the product path.

* ``grid``              acoustic_fdtd_bridge.py:291-317 (grid from the outline's
                        bounding box, padding, cell centres, back / plate layers)
* ``point_in_polygon``  acoustic_fdtd_bridge.py:424-453 (even-odd ray casting,
                        float64, same expression order)
* ``voxelise``          acoustic_fdtd_bridge.py:312-359 (cell types, plate mask,
                        exterior mask), written with array masks instead of
                        the reference's per-column loops
* ``face_fractions``    acoustic_fdtd_bridge.py:456-552 (cut-cell open fractions)

Pinned (tests/test_geometry_cpu.py) by analytic membership for convex and
L-shaped polygons, a hand-computed grid with binary-exact numbers, and
brute-force per-cell / per-face loops with analytic rectangle membership.
"""
from __future__ import annotations

import math

import numpy as np

AIR, WALL, PLATE = 0, 1, 2


def grid(poly, body_h, dx, pad_cells):
    poly = np.asarray(poly, np.float32)
    lo_x, hi_x = float(poly[:, 0].min()), float(poly[:, 0].max())
    lo_y, hi_y = float(poly[:, 1].min()), float(poly[:, 1].max())
    pad = pad_cells * dx
    org = (lo_x - pad, lo_y - pad, 0.0 - pad)
    top = (hi_x + pad, hi_y + pad, body_h + pad)
    dims = [max(4, int(math.ceil((b - a) / dx))) for a, b in zip(org, top)]
    centres = [o + (np.arange(n) + 0.5) * dx for o, n in zip(org, dims)]
    zs = centres[2]
    return dict(Nx=dims[0], Ny=dims[1], Nz=dims[2], dx=dx, gx_min=org[0], gy_min=org[1],
                gz_min=org[2], xs=centres[0], ys=centres[1], zs=zs,
                iz_back=int(np.argmin(np.abs(zs - 0.0))), plate_iz=int(np.argmin(np.abs(zs - body_h))),
                z_min=0.0, z_max=body_h)


def point_in_polygon(X, Y, poly):
    X = np.asarray(X, np.float64)
    Y = np.asarray(Y, np.float64)
    poly = np.asarray(poly, np.float32)
    odd = np.zeros(np.broadcast(X, Y).shape, bool)
    n = len(poly)
    for e in range(n):
        xa, ya = float(poly[e, 0]), float(poly[e, 1])
        xb, yb = float(poly[e - 1, 0]), float(poly[e - 1, 1])
        straddles = (ya > Y) != (yb > Y)
        with np.errstate(divide="ignore", invalid="ignore"):
            left = X < (xb - xa) * (Y - ya) / (yb - ya + 1e-15) + xa
        odd ^= straddles & left
    return odd


def voxelise(poly, body_h, dx=0.008, pad_cells=64):
    g = grid(poly, body_h, dx, pad_cells)
    Nx, Ny, Nz = g["Nx"], g["Ny"], g["Nz"]
    X, Y = np.meshgrid(g["xs"], g["ys"], indexing="ij")
    inside = point_in_polygon(X, Y, poly)
    band = (g["z_min"] - dx < g["zs"]) & (g["zs"] < g["z_max"] + dx)
    ct = np.where(~inside[:, :, None] & band[None, None, :], WALL, AIR).astype(np.uint8)
    ct[inside, g["iz_back"]] = WALL
    ct[inside, g["plate_iz"]] = PLATE
    cavity = np.zeros((Nx, Ny, Nz), bool)
    cavity[:, :, g["iz_back"] + 1:g["plate_iz"]] = inside[:, :, None]
    g["exterior_mask"] = ~cavity
    return ct, inside.astype(np.uint8), g


def face_fractions(ct, pa, poly, g):
    Nx, Ny, Nz = g["Nx"], g["Ny"], g["Nz"]
    dx, ox, oy, oz = g["dx"], g["gx_min"], g["gy_min"], g["gz_min"]
    zlo = oz + g["iz_back"] * dx
    zhi = oz + (g["plate_iz"] + 0.5) * dx
    solid = np.isin(np.asarray(ct).reshape(Nx, Ny, Nz), (WALL, PLATE))
    pa = np.asarray(pa).reshape(Nx, Ny).astype(bool)
    plate_rule = 0 < g["plate_iz"] < Nz - 1
    out = []
    for ax in range(3):
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[ax], hi[ax] = slice(0, -1), slice(1, None)
        sL, sR = solid[tuple(lo)], solid[tuple(hi)]
        frac = (~(sL | sR)).astype(np.float32)
        cut = sL ^ sR
        if ax == 2 and plate_rule:
            for kk in (g["plate_iz"] - 1, g["plate_iz"]):
                frac[:, :, kk][pa] = 1.0
                cut[:, :, kk] = False
        idx = np.nonzero(cut)
        if idx[0].size:
            # sub-point coordinates: the face's own axis at the face, the two
            # tangential axes at offsets 0.25 / 0.75 (first tangential outer)
            tang = [a for a in range(3) if a != ax]
            count = np.zeros(idx[0].size, np.float32)
            for oa in (0.25, 0.75):
                for ob in (0.25, 0.75):
                    pos = [None, None, None]
                    pos[ax] = (idx[ax] + 1) * dx
                    pos[tang[0]] = (idx[tang[0]] + oa) * dx
                    pos[tang[1]] = (idx[tang[1]] + ob) * dx
                    x, y, z = ox + pos[0], oy + pos[1], oz + pos[2]
                    hit = point_in_polygon(x, y, poly) & (z >= zlo) & (z <= zhi)
                    count += hit.astype(np.float32)
            frac[idx] = count / 4.0
        out.append(frac.ravel())
    return tuple(out)
