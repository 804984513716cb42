"""The two-step sweep of the reference's order-4 stencil (afdtd_fused4.hip)
against the float32 oracle, bit for bit: ragged grids (partial y and z
tiles, segments shorter than the sweep's five-plane pipeline), PML depths 0
to 10, point sources and receivers inside the ring of neighbouring tiles,
odd step counts (a trailing single step) and calls that cross the
stability-check step 256 (single steps there)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _pair(oracle_mod, Nx, Ny, Nz, n_pml, seg_x=0, seed=0):
    from afdtd import AcousticFDTD
    ct = np.zeros(Nx * Ny * Nz, np.uint8)
    pa = np.zeros(Nx * Ny, np.uint8)
    g = AcousticFDTD(Nx, Ny, Nz, 0.01, c=343.0, rho_air=1.21, cell_type=ct, plate_iz=Nz // 2,
                     plate_active=pa, n_pml=n_pml, kernel="fused2", seg_x=seg_x)
    assert g.kernel_family == "fused2"
    o = oracle_mod.OracleFDTD(Nx, Ny, Nz, 0.01, c=343.0, rho_air=1.21, cell_type=ct,
                              plate_iz=Nz // 2, plate_active=pa, n_pml=n_pml, precision="f32")
    assert g.dt == o.dt
    rng = np.random.default_rng(seed)
    P = rng.standard_normal(Nx * Ny * Nz).astype(np.float32)
    vx = (1e-3 * rng.standard_normal((Nx - 1) * Ny * Nz)).astype(np.float32)
    vy = (1e-3 * rng.standard_normal(Nx * (Ny - 1) * Nz)).astype(np.float32)
    vz = (1e-3 * rng.standard_normal(Nx * Ny * (Nz - 1))).astype(np.float32)
    g.set_state(P, vx, vy, vz)
    o.set_fields(P, vx, vy, vz)
    return g, o


def _same(g, o, what):
    Po, vxo, vyo, vzo = o.get_fields()
    for name, a, b in (("P", g.get_pressure_field().ravel(), Po), ("Vx", g.get_velocity_x(), vxo),
                       ("Vy", g.get_velocity_y(), vyo), ("Vz", g.get_velocity_z(), vzo)):
        nbad = int(np.count_nonzero(_bits(a) != _bits(b)))
        assert nbad == 0, f"{what} {name}: {nbad} of {a.size} values differ"


@pytest.mark.parametrize("shape,n_pml,seg_x", [
    ((21, 19, 70), 0, 0),
    ((21, 19, 70), 2, 0),
    ((30, 40, 131), 6, 7),
    ((14, 37, 66), 10, 3),
    ((9, 18, 64), 0, 1),
    ((40, 55, 129), 10, 0),
])
def test_order4_two_step_bitwise(oracle_mod, shape, n_pml, seg_x):
    Nx, Ny, Nz = shape
    print(f"fused4 {shape} n_pml={n_pml} seg_x={seg_x}", flush=True)
    g, o = _pair(oracle_mod, Nx, Ny, Nz, n_pml, seg_x, seed=Nx + Ny)
    for chunk in (2, 1, 6, 5):
        assert g.step(chunk) == 0 and o.step(chunk) == 0
        _same(g, o, f"{shape} pml={n_pml} seg={seg_x} after {g.step_count} steps")
    assert g.step_count == o.step_count == 14
    g.close()


def test_order4_two_step_sources_receivers_and_step_256(oracle_mod):
    """Sources on tile-ring cells (y = 17, 18, 36: rows of neighbouring
    18-row tiles; z = 63, 64: columns of neighbouring 64-column tiles) and
    receivers between them; 252 + 10 steps so that the pair schedule breaks
    at the stability-check step."""
    from afdtd import differentiated_gaussian
    Nx, Ny, Nz = 24, 40, 130
    g, o = _pair(oracle_mod, Nx, Ny, Nz, 4, seed=9)
    cells = [(10, 17, 63), (12, 18, 64), (11, 36, 3), (10, 17, 63)]
    src = [k + Nz * (j + Ny * i) for i, j, k in cells]
    sig = np.stack([differentiated_gaussian(262, 3.0, 8.0 + 2 * q, 1.0) for q in range(len(src))])
    rec = np.array([[10, 17, 63], [12.5, 18.25, 63.5], [11, 35.5, 2.75], [5, 5, 120],
                    [20.1, 38.9, 128.7]], np.float32)
    for m in (g, o):
        m.set_point_sources(src, sig)
        m.set_receivers(rec)
        m.record_receivers(262)
    for chunk in (252, 3, 7):
        assert g.step(chunk) == 0 and o.step(chunk) == 0
        _same(g, o, f"after {g.step_count} steps")
    tg, to = g.get_receiver_traces(), o.get_receiver_traces().astype(np.float32)
    assert tg.shape == (262, 5)
    assert np.array_equal(_bits(tg), _bits(to))
    assert np.abs(to).max() > 0
    g.close()
