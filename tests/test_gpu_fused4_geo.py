"""The two-step order-4 sweep with cell geometry (afdtd_fused4.hip, GEOM
build: walls, closed and fractional faces, an active plate whose second
update runs inside the sweep) against the float32 oracle and the single-step
geometry sweep, bit for bit.  Cases: ragged grids, PML 0 to 10, the plate's
two face columns inside a tile, on a tile seam (z = 63 | 64), one column
past it and on either z side of the grid, fractions 0, -0, 1 and in (0, 1) on open and plate faces, point
sources and receivers beside walls, and a plate guard that trips on the
second step of a pair (the pair's first step is kept)."""
import numpy as np
import pytest

from test_gpu_fused_geo import _fields, _same, _scene

pytestmark = pytest.mark.gpu


def _make(oracle_mod, shape, n_pml, seed, walls, plate, fracs, kernel, piz=None):
    from afdtd import AcousticFDTD
    Nx, Ny, Nz = shape
    ct, pa, piz0, fr = _scene(Nx, Ny, Nz, seed, walls, plate, fracs)
    piz = piz0 if piz is None else piz
    common = dict(c=343.0, rho_air=1.21, cell_type=ct.ravel(), plate_iz=piz,
                  plate_active=pa.ravel(), n_pml=n_pml)
    g = AcousticFDTD(Nx, Ny, Nz, 0.01, **common, order=4, kernel=kernel) if kernel else None
    o = oracle_mod.OracleFDTD(Nx, Ny, Nz, 0.01, **common, order=4, precision="f32")
    if fr is not None:
        if piz != piz0:        # the scene's plate-face fractions sit at piz0: move them
            fz = fr[2].reshape(Nx, Ny, Nz - 1)
            for (i, j, k), v in (((3, 4, piz - 1), 0.0), ((4, 5, piz), -0.0),
                                 ((5, 6, piz - 1), 0.5), ((6, 6, piz), 0.0)):
                if 0 <= k < Nz - 1:
                    fz[i, j, k] = v
        if g is not None:
            g.set_face_fractions(*fr)
        o.set_face_fractions(*fr)
    rng = np.random.default_rng(seed + 1)
    P = rng.standard_normal(Nx * Ny * Nz).astype(np.float32)
    octy = o.get_cell_type()
    P[(octy == 1) | (octy == 2)] = 0.0
    vx = (1e-3 * rng.standard_normal((Nx - 1) * Ny * Nz)).astype(np.float32)
    vy = (1e-3 * rng.standard_normal(Nx * (Ny - 1) * Nz)).astype(np.float32)
    vz = (1e-3 * rng.standard_normal(Nx * Ny * (Nz - 1))).astype(np.float32)
    if g is not None:
        g.set_state(P, vx, vy, vz)
    o.set_fields(P, vx, vy, vz)
    return g, o, piz


CASES = [
    # shape, n_pml, walls, plate, fracs, plate_iz (None: Nz // 2)
    ((20, 24, 16), 2, 0.0, True, False, None),
    ((21, 19, 70), 0, 0.06, True, True, None),
    ((17, 37, 66), 10, 0.04, True, True, None),
    ((23, 40, 131), 3, 0.03, False, True, None),
    ((19, 30, 128), 10, 0.02, True, True, 64),     # plate faces z = 63 | 64: a tile seam
    ((18, 27, 130), 4, 0.03, True, True, 65),      # faces 64, 65: next tile's first columns
    ((16, 50, 148), 10, 0.0, True, True, None),    # 148 = 64 + 64 + 20 columns
    ((14, 26, 40), 2, 0.02, True, True, 0),        # plate on the z = 0 side: no face below
    ((14, 26, 40), 0, 0.02, True, True, 39),       # on the far side: no face above
]


@pytest.mark.parametrize("shape,n_pml,walls,plate,fracs,piz", CASES)
def test_geometry_two_step_bitwise(oracle_mod, shape, n_pml, walls, plate, fracs, piz):
    print(f"geometry two-step {shape} pml={n_pml} walls={walls} plate={plate} fracs={fracs} "
          f"piz={piz}", flush=True)
    seed = 7 + shape[1]
    g, o, piz = _make(oracle_mod, shape, n_pml, seed, walls, plate, fracs, "fused2", piz)
    ref, _, _ = _make(oracle_mod, shape, n_pml, seed, walls, plate, fracs, "fused", piz)
    assert g.kernel_family == "fused" and g.steps_per_sweep == 2
    assert ref.steps_per_sweep == 1
    Nx, Ny, Nz = shape
    if plate:
        cells = np.array([piz + 1 + Nz * (j + Ny * i)
                          for i, j in ((3, 4), (5, 7), (Nx - 4, Ny - 3))], np.int32)
        w = np.array([1.0, 0.6, 0.3], np.float32)
        for m in (g, o, ref):
            m.set_bridge_sources(cells, w)
    for s in range(10):
        if plate:
            d = (2e-3 * np.array([np.sin(0.3 * s), np.cos(0.2 * s), 0.5], np.float32))
            for m in (g, o, ref):
                m.inject_bridge_drive(d, 1.0)
        n = (2, 3, 4, 1, 5)[s % 5]
        assert g.step(n) == 0 and o.step(n) == 0 and ref.step(n) == 0
        if s % 3 == 2 or s == 9:
            what = f"{shape} after {g.step_count} steps"
            for a, b, c, name in zip(_fields(g), _fields(o), _fields(ref), ("P", "Vx", "Vy", "Vz")):
                _same(a, b, f"{what} {name} vs oracle")
                _same(a, c, f"{what} {name} vs single-step sweep")
    assert g.step_count == o.step_count == 30
    if plate:
        _same(g.get_plate_displacement(), o.get_plate_displacement(), "plate displacement")
        assert np.abs(o.get_plate_displacement()).max() > 0
    g.close()
    ref.close()


def test_geometry_two_step_sources_receivers(oracle_mod):
    """Point sources and receivers beside walls and the plate, on tile-ring
    rows and columns, through pairs and a trailing single step."""
    from afdtd import differentiated_gaussian
    shape = (22, 40, 130)
    Nx, Ny, Nz = shape
    g, o, piz = _make(oracle_mod, shape, 4, 31, 0.02, True, True, "fused2")
    octy = o.get_cell_type().reshape(shape)
    cells = [(10, 23, 63), (12, 24, 64), (11, 36, 3), (9, 5, piz + 1), (10, 23, 63)]
    cells = [c for c in cells if octy[c] == 0]
    src = [k + Nz * (j + Ny * i) for i, j, k in cells]
    sig = np.stack([differentiated_gaussian(40, 3.0, 6.0 + 2 * q, 1.0) for q in range(len(src))])
    rec = np.array([[10, 23, 63], [12.5, 24.25, 63.5], [11, 35.5, 2.75], [9, 5, piz + 1.5],
                    [20.1, 38.9, 128.7]], np.float32)
    for m in (g, o):
        m.set_point_sources(src, sig)
        m.set_receivers(rec)
        m.record_receivers(40)
    for chunk in (8, 3, 29):
        assert g.step(chunk) == 0 and o.step(chunk) == 0
        for a, b, name in zip(_fields(g), _fields(o), ("P", "Vx", "Vy", "Vz")):
            _same(a, b, f"after {g.step_count} steps {name}")
    tg, to = g.get_receiver_traces(), o.get_receiver_traces().astype(np.float32)
    assert tg.shape == (40, 5)
    _same(tg, to, "receiver traces")
    assert np.abs(to).max() > 0
    g.close()


def test_geometry_two_step_plate_guard_on_second_step(oracle_mod):
    """A bridge kick sized so that the oracle's plate guard trips on the
    second of two steps: the pair keeps its first step (fields equal to the
    oracle's), reports the failure at that step, and continues as the
    single-step sweep does."""
    shape = (20, 24, 40)
    args = (shape, 2, 7, 0.0, True, True)
    Nx, Ny, Nz = shape

    def kick(m, piz, amp):
        cells = np.array([piz + 1 + Nz * (j + Ny * i) for i, j in ((8, 9), (9, 11))], np.int32)
        m.set_bridge_sources(cells, np.array([1.0, 0.7], np.float32))
        m.inject_bridge_drive(np.array([amp, amp], np.float32), 1.0)

    def kicked(amp, kernel):
        g, o, piz = _make(oracle_mod, *args, kernel)
        for m in (g, o):
            kick(m, piz, amp)
        return g, o

    amp, found = 1e4, None       # the kick moves the plate by about 2.4e-7 per unit
    for _ in range(30):
        _, o, piz = _make(oracle_mod, *args, None)
        kick(o, piz, amp)
        rc = o.step(2)
        if rc != 0:
            if o.step_count == 1:
                found = amp
            break
        amp *= 1.25
    assert found is not None, "no kick trips the guard on the second step"
    print(f"kick amplitude {found:.4g}", flush=True)
    g, o = kicked(found, "fused2")
    ref, _ = kicked(found, "fused")
    assert g.steps_per_sweep == 2
    assert o.step(2) != 0 and o.step_count == 1
    for m in (g, ref):
        with pytest.raises(RuntimeError, match="FDTD_ERR_UNSTABLE"):
            m.step(2)
        assert m.step_count == 1
    for a, b, c, name in zip(_fields(g), _fields(o), _fields(ref), ("P", "Vx", "Vy", "Vz")):
        _same(a, b, f"after the guard {name} vs oracle")
        _same(a, c, f"after the guard {name} vs single-step sweep")
    _same(g.get_plate_displacement(), ref.get_plate_displacement(), "plate after the guard")
    g.close()
    ref.close()
