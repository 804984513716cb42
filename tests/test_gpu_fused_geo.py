"""The single-step sweep with cell geometry (walls, plate cells, closed and
fractional faces, an active plate) against the float32 oracle and against
the generic kernels, bit for bit.

Cases: both stencil orders, PML depths 0 to 10, tile heights 8 / 16 / 32,
x-segments shorter than the sweep's pipeline, ragged grids (partial y and z
tiles); random walls, a plate layer with a bridge drive, and face fractions
that include the values whose divergence multiply can change a bit: 0 and
-0 on open and on plate faces, 1, and fractions in (0, 1).  The voxelised
two-bout body at the factory's spacing, scaled 4x (395 x 549 x 148 cells),
runs the bridge's per-sample loop against the oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _same(a, b, what):
    nbad = int(np.count_nonzero(_bits(a) != _bits(b)))
    assert nbad == 0, f"{what}: {nbad} of {np.size(a)} values differ"


def _fields(m):
    if hasattr(m, "get_fields"):
        return m.get_fields()
    return (m.get_pressure_field().ravel(), m.get_velocity_x(), m.get_velocity_y(),
            m.get_velocity_z())


def _scene(Nx, Ny, Nz, seed, walls, plate, fracs):
    rng = np.random.default_rng(seed)
    ct = np.zeros((Nx, Ny, Nz), np.uint8)
    if walls:
        ct[rng.random((Nx, Ny, Nz)) < walls] = 1
        ct[2:5, 3:9, 1:Nz - 2] = 1                       # a solid block
    piz = Nz // 2
    pa = np.zeros((Nx, Ny), np.uint8)
    if plate:
        pa[1:Nx - 2, 2:Ny - 1] = 1
        pa[Nx // 2, Ny // 2] = 0                         # a hole in the plate
    fr = None
    if fracs:
        nvx, nvy, nvz = (Nx - 1) * Ny * Nz, Nx * (Ny - 1) * Nz, Nx * Ny * (Nz - 1)
        fr = []
        for n in (nvx, nvy, nvz):
            u = rng.random(n)
            f = np.ones(n, np.float32)
            f[u < 0.15] = rng.uniform(0.05, 0.95, int((u < 0.15).sum()))
            f[(u >= 0.15) & (u < 0.2)] = 0.0
            f[(u >= 0.2) & (u < 0.23)] = -0.0
            fr.append(f)
        # the plate faces (k = piz-1, piz in the packed z layout): 0, -0 and a fraction
        fz = fr[2].reshape(Nx, Ny, Nz - 1)
        fz[3, 4, piz - 1] = 0.0
        fz[4, 5, piz] = -0.0
        fz[5, 6, piz - 1] = 0.5
        fz[6, 6, piz] = 0.0
    return ct, pa, piz, fr


def _build(oracle_mod, shape, order, n_pml, seed, walls, plate, fracs, kernel, tile_y, seg_x):
    from afdtd import AcousticFDTD
    Nx, Ny, Nz = shape
    ct, pa, piz, fr = _scene(Nx, Ny, Nz, seed, walls, plate, fracs)
    common = dict(c=343.0, rho_air=1.21, cell_type=ct.ravel(), plate_iz=piz,
                  plate_active=pa.ravel(), n_pml=n_pml)
    g = AcousticFDTD(Nx, Ny, Nz, 0.01, **common, order=order, kernel=kernel, tile_y=tile_y,
                     seg_x=seg_x)
    o = oracle_mod.OracleFDTD(Nx, Ny, Nz, 0.01, **common, order=order, precision="f32")
    assert g.dt == o.dt
    if fr is not None:
        g.set_face_fractions(*fr)
        o.set_face_fractions(*fr)
    assert g.kernel_family == ("generic" if kernel == "generic" else "fused")
    rng = np.random.default_rng(seed + 1)
    P = rng.standard_normal(Nx * Ny * Nz).astype(np.float32)
    octy = o.get_cell_type()
    P[(octy == 1) | (octy == 2)] = 0.0
    vx = (1e-3 * rng.standard_normal((Nx - 1) * Ny * Nz)).astype(np.float32)
    vy = (1e-3 * rng.standard_normal(Nx * (Ny - 1) * Nz)).astype(np.float32)
    vz = (1e-3 * rng.standard_normal(Nx * Ny * (Nz - 1))).astype(np.float32)
    g.set_state(P, vx, vy, vz)
    o.set_fields(P, vx, vy, vz)
    if plate:
        cells = np.array([piz + 1 + Nz * (j + Ny * i) for i, j in ((3, 4), (5, 7), (Nx - 4, Ny - 3))],
                         np.int32)
        w = np.array([1.0, 0.6, 0.3], np.float32)
        g.set_bridge_sources(cells, w)
        o.set_bridge_sources(cells, w)
    return g, o


CASES = [
    # shape, order, n_pml, tile_y, seg_x, walls, plate, fracs
    ((20, 24, 16), 4, 2, 16, 0, 0.0, True, False),
    ((21, 19, 70), 4, 0, 16, 0, 0.06, True, True),
    ((17, 37, 66), 4, 10, 8, 3, 0.04, True, True),
    ((23, 40, 131), 4, 3, 32, 7, 0.03, False, True),
    ((19, 18, 64), 2, 0, 16, 1, 0.08, True, True),
    ((22, 33, 90), 2, 4, 8, 0, 0.05, False, False),
    ((18, 35, 129), 2, 6, 32, 5, 0.02, True, True),
]


@pytest.mark.parametrize("shape,order,n_pml,tile_y,seg_x,walls,plate,fracs", CASES)
def test_geometry_sweep_bitwise(oracle_mod, shape, order, n_pml, tile_y, seg_x, walls, plate,
                                fracs):
    print(f"geometry sweep {shape} order={order} pml={n_pml} tile_y={tile_y} seg_x={seg_x} "
          f"walls={walls} plate={plate} fracs={fracs}", flush=True)
    g, o = _build(oracle_mod, shape, order, n_pml, 11 + shape[1], walls, plate, fracs, "fused",
                  tile_y, seg_x)
    ref, _ = _build(oracle_mod, shape, order, n_pml, 11 + shape[1], walls, plate, fracs,
                    "generic", 0, 0)
    for s in range(12):
        if plate:
            d = (2e-3 * np.array([np.sin(0.3 * s), np.cos(0.2 * s), 0.5], np.float32))
            for m in (g, o, ref):
                m.inject_bridge_drive(d, 1.0)
        n = 1 + s % 3
        assert g.step(n) == 0 and o.step(n) == 0 and ref.step(n) == 0
        if s % 4 == 3:
            what = f"{shape} order {order} after {g.step_count} steps"
            for a, b, c, name in zip(_fields(g), _fields(o), _fields(ref), ("P", "Vx", "Vy", "Vz")):
                _same(a, b, f"{what} {name} vs oracle")
                _same(a, c, f"{what} {name} vs generic")
    if plate:
        _same(g.get_plate_displacement(), o.get_plate_displacement(), "plate displacement")
        assert np.abs(o.get_plate_displacement()).max() > 0
    g.close()
    ref.close()


def test_fractions_switch_two_step_handle_to_geometry_sweep(oracle_mod):
    """A room on the two-step sweep that receives face fractions continues on
    the single-step sweep with geometry, from the same state."""
    from afdtd import AcousticFDTD
    Nx, Ny, Nz = 20, 26, 70
    ct = np.zeros(Nx * Ny * Nz, np.uint8)
    pa = np.zeros(Nx * Ny, np.uint8)
    kw = dict(c=343.0, rho_air=1.21, cell_type=ct, plate_iz=Nz // 2, plate_active=pa, n_pml=3)
    g = AcousticFDTD(Nx, Ny, Nz, 0.01, **kw)
    o = oracle_mod.OracleFDTD(Nx, Ny, Nz, 0.01, **kw, precision="f32")
    assert g.kernel_family == "fused2"
    rng = np.random.default_rng(5)
    P = rng.standard_normal(Nx * Ny * Nz).astype(np.float32)
    g.set_state(P, *[np.zeros(n, np.float32) for n in g.get_velocity_dims()])
    o.set_fields(P, *[np.zeros(n, np.float32) for n in g.get_velocity_dims()])
    assert g.step(6) == 0 and o.step(6) == 0
    fr = [np.where(rng.random(n) < 0.2, rng.uniform(0, 1, n), 1.0).astype(np.float32)
          for n in g.get_velocity_dims()]
    g.set_face_fractions(*fr)
    o.set_face_fractions(*fr)
    assert g.kernel_family == "fused"
    assert g.step(9) == 0 and o.step(9) == 0
    for a, b, name in zip(_fields(g), _fields(o), ("P", "Vx", "Vy", "Vz")):
        _same(a, b, f"after the switch: {name}")
    g.close()


def test_body_x4_bridge_loop_matches_oracle(oracle_mod):
    """The bench's body (two-bout outline x4, dx 4 mm, soundhole, 24 cells of
    padding, PML 10, cut-cell fractions, active plate, three Gaussian bridge
    sources) through the bridge's per-sample loop: 4 audio samples of random
    drives, fields and plate equal to the oracle's after each block."""
    from afdtd.bridge import FDTDCoEvolver
    from afdtd.geometry import apply_soundhole_cutout, build_bridge_kernel, two_bout_outline
    from oracle import geometry as og
    scale, dx, pad = 4.0, 0.004, 24
    poly = (two_bout_outline(256) * scale).astype(np.float32)
    body_h = 0.1 * scale
    bridges = [(-0.08, -0.52), (0.0, -0.52), (0.08, -0.52)]
    hole = (0.0, 0.24, 0.18)
    ev = FDTDCoEvolver.from_outline(poly, body_h, bridges, soundhole=hole, dx=dx,
                                    pad_cells=pad, n_pml=10)
    info = ev.info
    Nx, Ny, Nz = info["Nx"], info["Ny"], info["Nz"]
    print(f"body {Nx}x{Ny}x{Nz}, kernel {ev.fdtd.kernel_family}, "
          f"{ev.steps_per_sample} steps per sample", flush=True)
    assert ev.fdtd.kernel_family == "fused"
    ct, pa, g = og.voxelise(poly, body_h, dx, pad)
    apply_soundhole_cutout(ct, pa, g, hole)
    o = oracle_mod.OracleFDTD(Nx, Ny, Nz, dx, c=343.0, rho_air=1.21, cell_type=ct.ravel(),
                              plate_iz=g["plate_iz"], plate_active=pa.ravel(), n_pml=10,
                              precision="f32")
    o.set_face_fractions(*og.face_fractions(ct, pa, poly, g))
    idx, wgt = build_bridge_kernel(bridges, info)
    o.set_bridge_sources(idx, wgt)
    rng = np.random.default_rng(77)
    drives = (1e-3 * rng.standard_normal((4, len(idx)))).astype(np.float32)
    for blk in (drives[:1], drives[1:]):
        ev.drive_bridge_block(blk)
        for row in blk:
            o.inject_bridge_drive(row, 1.0)
            assert o.step(ev.steps_per_sample) == 0
        for a, b, name in zip(_fields(ev.fdtd), o.get_fields(), ("P", "Vx", "Vy", "Vz")):
            _same(a, b, f"body after {o.step_count} steps: {name}")
        _same(ev.get_plate_displacement(), o.get_plate_displacement(), "body plate")
    assert np.abs(o.get_plate_displacement()).max() > 0
